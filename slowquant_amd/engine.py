"""Host objects over the HIP engine: CI space, gate programs, Hamiltonians.

Device memory and streams come from PyTorch (plumbing only); every numerical
operation runs in libsqhip.so.  State vectors are float64 CUDA tensors in the
engine's internal (alpha block, beta block) sign convention; use
``Space.from_reference`` / ``Space.to_reference`` at the boundary with
reference-convention ``ci_coeffs``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError("engine buffers must be contiguous float64 CUDA tensors")
    return t.data_ptr()


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("the HIP state-vector engine needs a GPU (torch.cuda.is_available() is False)")
    N.lib()


class Space:
    """CI space of (n_orb, n_alpha, n_beta) built on the device (ci_spaces.get_indexing)."""

    def __init__(self, n_orb: int, n_alpha: int, n_beta: int) -> None:
        require_gpu()
        self.n_orb, self.n_alpha, self.n_beta = n_orb, n_alpha, n_beta
        h = ctypes.c_void_p()
        N.check(N.lib().sq_space_create(n_orb, n_alpha, n_beta, ctypes.byref(h)))
        self._h = h
        self.dim = int(N.lib().sq_space_dim(h))
        na, nb = ctypes.c_int32(), ctypes.c_int32()
        N.check(N.lib().sq_space_nstrings(h, ctypes.byref(na), ctypes.byref(nb)))
        self.n_alpha_strings, self.n_beta_strings = na.value, nb.value
        self._work: dict[int, torch.Tensor] = {}

    @property
    def handle(self):
        return self._h

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        self._h = None
        try:
            if h is not None and N._lib is not None:
                N.lib().sq_space_destroy(h)
        except (AttributeError, TypeError):  # interpreter shutdown: module globals already cleared
            pass

    def work(self, what: int) -> torch.Tensor:
        if what not in self._work:
            nbytes = int(N.lib().sq_work_bytes(self._h, what))
            if nbytes < 0:
                raise ValueError("unknown workspace kind")
            self._work[what] = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device="cuda")
        return self._work[what]

    def configure_sigma(self, rows_per_block: int = 0, nsplit: int = 0, deterministic: int = -1) -> None:
        """Row-sigma kernel shape on this space (0: automatic); deterministic=1
        selects fixed-order sums, 0 LDS-atomic sums, -1 the default."""
        N.check(N.lib().sq_sigma_configure(self._h, int(rows_per_block), int(nsplit), int(deterministic)))

    def strings(self) -> tuple[np.ndarray, np.ndarray]:
        a = np.zeros(self.n_alpha_strings, dtype=np.uint32)
        b = np.zeros(self.n_beta_strings, dtype=np.uint32)
        N.check(N.lib().sq_space_copy_strings(self._h, a.ctypes.data, b.ctypes.data))
        return a, b

    def idx2det(self) -> np.ndarray:
        out = np.zeros(self.dim, dtype=np.int64)
        N.check(N.lib().sq_space_idx2det(self._h, out.ctypes.data, _stream()))
        return out

    def from_reference(self, ref) -> torch.Tensor:
        x = torch.as_tensor(np.ascontiguousarray(ref, dtype=np.float64) if not torch.is_tensor(ref) else ref)
        x = x.to(device="cuda", dtype=torch.float64).contiguous()
        out = torch.empty_like(x)
        N.check(N.lib().sq_state_from_reference(self._h, _ptr(x), _ptr(out), _stream()))
        return out

    def to_reference(self, internal: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(internal)
        N.check(N.lib().sq_state_to_reference(self._h, _ptr(internal), _ptr(out), _stream()))
        return out

    def to_reference_numpy(self, internal: torch.Tensor) -> np.ndarray:
        return self.to_reference(internal).cpu().numpy()


class Program:
    """A gate list (UpsStructure / UccStructure layout) compiled for a space."""

    def __init__(self, space: Space, types, indices) -> None:
        self.space = space
        self.types = list(types)
        self.indices = [tuple(int(x) for x in t) for t in indices]
        descs = N.gate_descs(self.types, self.indices)
        h = ctypes.c_void_p()
        N.check(N.lib().sq_program_create(space.handle, descs, len(self.types), ctypes.byref(h)))
        self._h = h
        self.n_gates = len(self.types)

    @property
    def handle(self):
        return self._h

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        self._h = None
        try:
            if h is not None and N._lib is not None:
                N.lib().sq_program_destroy(h)
        except (AttributeError, TypeError):  # interpreter shutdown: module globals already cleared
            pass

    def set_tile_shape(self, target: int = 0, threads: int = 0) -> None:
        """Elements per tile (0: automatic) and threads per tile block (256 or
        512; 0: 512) of the fused rotation stages."""
        N.check(N.lib().sq_program_set_tile_shape(self._h, int(target), int(threads)))

    def theta_tensor(self, theta) -> torch.Tensor:
        if isinstance(theta, torch.Tensor):
            t = theta.detach().to(device="cuda", dtype=torch.float64)
            if t.numel() != self.n_gates:
                raise ValueError(f"Expected {self.n_gates} theta values got {t.numel()}")
            return t.contiguous()
        arr = np.asarray(theta, dtype=np.float64).ravel()
        if arr.size != self.n_gates:
            raise ValueError(f"Expected {self.n_gates} theta values got {arr.size}")
        # host values go through one pinned staging buffer (asynchronous DMA on
        # the engine's stream); the previous upload must have left it first
        if getattr(self, "_th_host", None) is None:
            n = max(self.n_gates, 1)
            self._th_host = torch.empty(n, dtype=torch.float64, pin_memory=True)
            self._th_dev = torch.empty(n, dtype=torch.float64, device="cuda")
            self._th_done = torch.cuda.Event()
            self._th_done.record()
        self._th_done.synchronize()
        self._th_host.numpy()[: arr.size] = arr
        dev = self._th_dev[: arr.size]
        dev.copy_(self._th_host[: arr.size], non_blocking=True)
        self._th_done.record()
        return dev

    def state_prep(self, theta, psi_in: torch.Tensor, dagger: bool = False, out: torch.Tensor | None = None):
        th = self.theta_tensor(theta)
        if out is None:
            out = torch.empty_like(psi_in)
        N.check(
            N.lib().sq_state_prep(
                self.space.handle, self._h, _ptr(th), _ptr(psi_in), _ptr(out), int(dagger), _ptr(self.space.work(0)), _stream()
            )
        )
        return out

    def apply_gate(self, gate: int, theta, psi: torch.Tensor) -> torch.Tensor:
        th = self.theta_tensor(theta)
        N.check(N.lib().sq_apply_gate(self.space.handle, self._h, gate, _ptr(th), _ptr(psi), _ptr(self.space.work(0)), _stream()))
        return psi

    def grad_action(self, gate: int, psi: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(psi)
        N.check(N.lib().sq_grad_action(self.space.handle, self._h, gate, _ptr(psi), _ptr(out), _stream()))
        return out

    def energy_grad(self, H: "Hamiltonian", theta, psi0: torch.Tensor, method: str = "adjoint"):
        th = self.theta_tensor(theta)
        e = ctypes.c_double()
        g = np.zeros(max(self.n_gates, 1))
        fn = {
            "adjoint": N.lib().sq_energy_grad_adjoint,
            "plain": N.lib().sq_energy_grad_plain,
            "reference": N.lib().sq_energy_grad_reference_sweep,
        }[method]
        N.check(
            fn(self.space.handle, self._h, H.handle, _ptr(th), _ptr(psi0), ctypes.byref(e), g.ctypes.data,
               _ptr(self.space.work(2)), _stream())
        )
        return e.value, g[: self.n_gates].copy()


    def adjoint_reverse(self, theta, psi: torch.Tensor, lam: torch.Tensor) -> np.ndarray:
        """dE/dtheta from psi = U(theta) psi0 and lam = H psi; both are consumed."""
        th = self.theta_tensor(theta)
        g = np.zeros(max(self.n_gates, 1))
        N.check(N.lib().sq_adjoint_reverse(self.space.handle, self._h, _ptr(th), _ptr(psi), _ptr(lam), g.ctypes.data,
                                           _ptr(self.space.work(2)), _stream()))
        return g[: self.n_gates].copy()


class Hamiltonian:
    """Active-space Hamiltonian e_core + h.E + 1/2 g.e on the device."""

    def __init__(self, space: Space, h: np.ndarray, g: np.ndarray, e_core: float = 0.0) -> None:
        n = space.n_orb
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(n, n)
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(n, n, n, n)
        self.space = space
        self.e_core = float(e_core)
        hh = ctypes.c_void_p()
        N.check(N.lib().sq_hamiltonian_create(space.handle, h.ctypes.data, g.ctypes.data, self.e_core, ctypes.byref(hh)))
        self._h = hh

    @property
    def handle(self):
        return self._h

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        self._h = None
        try:
            if h is not None and N._lib is not None:
                N.lib().sq_hamiltonian_destroy(h)
        except (AttributeError, TypeError):  # interpreter shutdown: module globals already cleared
            pass

    def sigma(self, psi: torch.Tensor, want_energy: bool = True):
        out = torch.empty_like(psi)
        e = ctypes.c_double()
        N.check(
            N.lib().sq_sigma(self.space.handle, self._h, _ptr(psi), _ptr(out), ctypes.byref(e) if want_energy else None,
                             _ptr(self.space.work(1)), _stream())
        )
        return out, (e.value if want_energy else None)


def rdm12(space: Space, psi: torch.Tensor) -> tuple[np.ndarray, np.ndarray]:
    n = space.n_orb
    r1 = np.zeros((n, n))
    r2 = np.zeros((n, n, n, n))
    N.check(N.lib().sq_rdm12(space.handle, _ptr(psi), r1.ctypes.data, r2.ctypes.data, _ptr(space.work(3)), _stream()))
    return r1, r2


def hf_state(space: Space, occ_det_index: int = 0) -> torch.Tensor:
    """Reference-convention unit vector at ``occ_det_index`` converted to internal."""
    ref = torch.zeros(space.dim, dtype=torch.float64, device="cuda")
    ref[occ_det_index] = 1.0
    return space.from_reference(ref)


class Operator:
    """A FermionicOperator (active-space, already folded) compiled for a space."""

    def __init__(self, space: Space, items, drop_outside: bool = False) -> None:
        items = list(items)
        arr = (N.OpString * max(len(items), 1))()
        for k, (key, coef) in enumerate(items):
            if len(key) > 24:
                raise ValueError("operator string longer than 24 operators")
            arr[k].coef = float(coef)
            arr[k].n_ops = len(key)
            for j, (idx, dag) in enumerate(key):
                arr[k].spin_orb[j] = int(idx)
                arr[k].is_creator[j] = 1 if dag else 0
        h = ctypes.c_void_p()
        N.check(N.lib().sq_operator_create(space.handle, arr, len(items), int(drop_outside), ctypes.byref(h)))
        self._h = h
        self.space = space

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        self._h = None
        try:
            if h is not None and N._lib is not None:
                N.lib().sq_operator_destroy(h)
        except (AttributeError, TypeError):  # interpreter shutdown: module globals already cleared
            pass

    def apply(self, psi: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(psi)
        N.check(N.lib().sq_operator_apply(self.space.handle, self._h, _ptr(psi), _ptr(out), _stream()))
        return out

    def matrix(self, reference_convention: bool = True) -> torch.Tensor:
        """Dense <I|O|J> on the device (dim x dim), reference sign convention by default."""
        d = self.space.dim
        M = torch.empty((d, d), dtype=torch.float64, device="cuda")
        N.check(N.lib().sq_operator_matrix(self.space.handle, self._h, int(reference_convention), _ptr(M), _stream()))
        return M


_SPACES: dict = {}


def space_for(n_orb: int, n_alpha: int, n_beta: int) -> Space:
    """Process-wide cache of device CI spaces."""
    key = (n_orb, n_alpha, n_beta)
    if key not in _SPACES:
        _SPACES[key] = Space(*key)
    return _SPACES[key]


def _ucc_work(space: Space, prog: Program) -> torch.Tensor:
    nbytes = int(N.lib().sq_ucc_work_bytes(space.handle, prog.handle))
    key = ("ucc", nbytes)
    if key not in space._work:
        space._work[key] = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device="cuda")
    return space._work[key]


def ucc_expm(prog: Program, theta, psi_in: torch.Tensor, t: float = 1.0) -> torch.Tensor:
    """exp(t T(theta)) psi_in for a UCC program (new tensor)."""
    th = prog.theta_tensor(theta)
    out = torch.empty_like(psi_in)
    N.check(N.lib().sq_ucc_expm(prog.space.handle, prog.handle, _ptr(th), float(t), _ptr(psi_in), _ptr(out),
                                _ptr(_ucc_work(prog.space, prog)), _stream()))
    return out


def ucc_energy_grad(prog: Program, H: Hamiltonian, theta, psi0: torch.Tensor, want_grad: bool = True):
    """E = <psi|H|psi>, psi = exp(T(theta)) psi0, and the analytic dE/dtheta."""
    th = prog.theta_tensor(theta)
    e = ctypes.c_double()
    g = np.zeros(max(prog.n_gates, 1))
    N.check(N.lib().sq_ucc_energy_grad(prog.space.handle, prog.handle, H.handle, _ptr(th), _ptr(psi0),
                                       ctypes.byref(e), g.ctypes.data if want_grad else None,
                                       _ptr(_ucc_work(prog.space, prog)), _stream()))
    return e.value, (g[: prog.n_gates].copy() if want_grad else None)

class Batch:
    """Many independent instances (space, rotation program, Hamiltonian) whose
    energies and adjoint gradients come from one persistent launch, started
    from each space's HF determinant or from given states (set_start_states).  ``run(theta, graph=True)`` replays a
    captured hipGraph of that launch."""

    def __init__(self, instances, n_workgroups: int = 0) -> None:
        self.instances = list(instances)
        n = len(self.instances)
        sp = (ctypes.c_void_p * n)(*[s.handle.value for s, _, _ in self.instances])
        pr = (ctypes.c_void_p * n)(*[p.handle.value for _, p, _ in self.instances])
        hh = (ctypes.c_void_p * n)(*[h.handle.value for _, _, h in self.instances])
        b = ctypes.c_void_p()
        N.check(N.lib().sq_batch_create(n, sp, pr, hh, int(n_workgroups), ctypes.byref(b)))
        self._h = b
        self.n_workgroups = int(n_workgroups)
        self.shape = (2, 1)  # the library's default kernel shape
        self.tile_shape = None  # the grouped programs' tile shape when an autotuner chose it
        self.n_params = int(N.lib().sq_batch_num_params(b))
        self.offsets = np.cumsum([0] + [p.n_gates for _, p, _ in self.instances])

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        self._h = None
        try:
            if h is not None and N._lib is not None:
                N.lib().sq_batch_destroy(h)
        except (AttributeError, TypeError):  # interpreter shutdown: module globals already cleared
            pass

    def set_theta(self, theta) -> None:
        th = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.float64).ravel() for t in theta])
                                  if isinstance(theta, (list, tuple)) else theta, dtype=np.float64)
        if th.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {th.size}")
        self._theta_host = th  # kept alive until the asynchronous copy has run
        N.check(N.lib().sq_batch_set_theta(self._h, th.ctypes.data, _stream()))

    def set_start_states(self, states) -> None:
        """Start state per instance: a CUDA float64 tensor of the instance's
        space dimension (engine convention, e.g. Space.from_reference) or None
        for the HF determinant.  The tensors are kept referenced here, since
        every later launch reads them."""
        states = list(states)
        if len(states) != len(self.instances):
            raise ValueError(f"expected {len(self.instances)} start states, got {len(states)}")
        kept, ptrs = [], []
        for (sp, _, _), x in zip(self.instances, states):
            if x is None:
                ptrs.append(None)
                continue
            if not (torch.is_tensor(x) and x.is_cuda and x.dtype == torch.float64 and x.is_contiguous()):
                raise TypeError("start states must be contiguous float64 CUDA tensors")
            if x.numel() != sp.dim:
                raise ValueError(f"start state has {x.numel()} amplitudes, its space has {sp.dim}")
            kept.append(x)
            ptrs.append(x.data_ptr())
        arr = (ctypes.c_void_p * len(ptrs))(*ptrs)
        N.check(N.lib().sq_batch_set_start_states(self._h, arr, _stream()))
        self._starts = kept

    def configure(self, rot_pairs: int = 2, gemm_tiles: int = 1) -> None:
        """Kernel shape: rotation pairs per thread in flight (1, 2, 4) and sigma
        GEMM column tiles per wave (1, 2)."""
        N.check(N.lib().sq_batch_configure(self._h, int(rot_pairs), int(gemm_tiles)))
        self.shape = (int(rot_pairs), int(gemm_tiles))

    def launch(self, graph: bool = False) -> None:
        fn = N.lib().sq_batch_run_graph if graph else N.lib().sq_batch_run
        N.check(fn(self._h, _stream()))

    def results(self) -> tuple[np.ndarray, list[np.ndarray]]:
        e = np.zeros(len(self.instances))
        g = np.zeros(max(self.n_params, 1))
        N.check(N.lib().sq_batch_results(self._h, e.ctypes.data, g.ctypes.data, _stream()))
        return e, [g[self.offsets[i]:self.offsets[i + 1]].copy() for i in range(len(self.instances))]

    def run(self, theta, graph: bool = False) -> tuple[np.ndarray, list[np.ndarray]]:
        self.set_theta(theta)
        self.launch(graph)
        return self.results()


SHAPES = tuple((u, ct) for u in (1, 2, 4) for ct in (1, 2))


# Instances of this many determinants or more run in the library's batched
# groups (sq_batch.hip, kGroupMinDim): through the tile-stage kernels, whose
# tile shape (elements per tile, threads per block) is the program's.
BATCH_GROUP_MIN_DIM = 16384
GROUP_TILE_SHAPES = ((0, 0), (1024, 256), (2048, 256), (4096, 512))


def _graph_ms(b: "Batch", reps: int) -> float:
    b.launch(graph=True)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.launch(graph=True)
    t1.record()
    t1.synchronize()
    return t0.elapsed_time(t1)


def grouped_programs(instances) -> list:
    """The distinct programs of the instances the batch runs in groups."""
    out, seen = [], set()
    for s, p, _ in instances:
        if s.dim >= BATCH_GROUP_MIN_DIM and id(p) not in seen:
            seen.add(id(p))
            out.append(p)
    return out


def autotune_batch(instances, theta, candidates=(64, 128, 256, 512, 1024), shapes=SHAPES, reps: int = 3,
                   tile_shapes=GROUP_TILE_SHAPES) -> Batch:
    """First, when some instances run in groups, time graph replays of a
    batch per tile shape in ``tile_shapes`` (set on the grouped programs)
    and keep the fastest.  Then build one Batch per workgroup count in
    ``candidates`` (capped at the number of instances); for each, time
    ``reps`` graph replays of every kernel shape in ``shapes`` ((rotation
    pairs in flight, GEMM column tiles per wave)) and return the fastest
    batch, configured with its best shape (its ``tile_shape`` attribute holds
    the grouped programs' tile shape, None without groups)."""
    big = grouped_programs(instances)
    best_ts = None
    if big:
        best_tt = float("inf")
        for ts in tile_shapes:
            for p in big:
                p.set_tile_shape(*ts)
            b = Batch(instances, 256)
            b.set_theta(theta)
            b.configure(4, 1)
            t = _graph_ms(b, reps)
            del b
            if t < best_tt:
                best_tt, best_ts = t, tuple(ts)
        for p in big:
            p.set_tile_shape(*best_ts)
    best, best_t, best_shape = None, float("inf"), None
    seen = set()
    for c in candidates:
        c = min(int(c), len(instances))
        if c in seen:
            continue
        seen.add(c)
        b = Batch(instances, c)
        b.set_theta(theta)
        b_best_t, b_shape = float("inf"), None
        for shape in shapes:
            b.configure(*shape)
            t = _graph_ms(b, reps)
            if t < b_best_t:
                b_best_t, b_shape = t, shape
        if b_best_t < best_t:
            best, best_t, best_shape = b, b_best_t, b_shape
    best.configure(*best_shape)
    best.tuned_ms = best_t / reps
    best.tile_shape = best_ts
    return best

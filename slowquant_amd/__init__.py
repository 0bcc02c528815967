"""MI355X-native state-vector engine for SlowQuant's UPS and UCC wave functions."""

"""Seeded synthetic workloads: OHLCV data and factor graphs used by the
benchmark, the smoke check and the tests."""

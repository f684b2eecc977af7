"""Support code: seeded synthetic phantoms for the benchmark and test workloads."""

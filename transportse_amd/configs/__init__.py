"""TRANSPORTSE config modules for the benchmark workloads (importable by
path with ``transportse_amd.load_config`` or ``python -m
transportse_amd.driver <file>``)."""

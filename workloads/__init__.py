"""Benchmark and test workloads: seeded synthetic references and read sets, and
a host restatement of the reference's build_cpu + to_bytes used to cross-check
the device index builder.  Not part of the query path (libhsfm.so)."""

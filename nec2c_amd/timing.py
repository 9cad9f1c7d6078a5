"""Timing documents of GPU solves and of the serving loop.

The reference writes a timing artifact with per-phase seconds and calls
(performance.c:150-225, schema "nec2c-timing" version 2).  This module writes
the same document for work done by libvie_hip.so, as schema version 3:
version 2 with "hip" added to the fft_backend values and an optional
"serving" object (throughput, p50/p99 latency, cache and admission figures of
nec2c_amd/serve.py).  Its JSON Schema is nec2c_amd/schemas/nec2c-timing-v3.schema.json.
"""
from __future__ import annotations

import json
import os
import resource
import time

SCHEMA_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schemas", "nec2c-timing-v3.schema.json")
PRODUCER = "nec2c_amd"
VERSION = "1"

SERVING_KEYS = ("requests", "rejected", "solves", "seconds", "solves_per_s", "requests_per_s", "p50_ms", "p99_ms",
                "service_p50_ms", "service_p99_ms", "cache_hits", "system_builds", "system_build_seconds",
                "converged", "max_relative_residual", "iterations_total")


def document(deck, phases, wall_seconds, cpu_seconds=None, success=True, fft_requested="hip", fft_used="hip",
             serving=None):
    """A timing document: ``phases`` is a list of {"name", "seconds", "calls"}
    (vie.phases(); phases without calls are left out, as the reference never
    writes one), ``serving`` a serve.Server report (optional)."""
    if cpu_seconds is None:
        cpu_seconds = time.process_time()
    rss = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
    doc = {
        "schema": "nec2c-timing",
        "schema_version": 3,
        "producer": {"program": PRODUCER, "program_version": VERSION},
        "input": {"deck": str(deck)},
        "dense_backend": {"requested": None, "used": None},
        "fft_backend": {"requested": fft_requested, "used": fft_used},
        "concurrency": None,
        "clock": {"wall": "monotonic", "cpu": "process"},
        "completion": {"status": "ok" if success else "error"},
        "process": {"wall_seconds": max(0.0, float(wall_seconds)), "cpu_seconds": max(0.0, float(cpu_seconds)),
                    "peak_rss_kib": int(rss) if rss >= 0 else None},
        "phases": [{"name": p["name"], "seconds": max(0.0, float(p["seconds"])), "calls": int(p["calls"])}
                   for p in phases if int(p["calls"]) >= 1],
    }
    if serving is not None:
        doc["serving"] = {k: serving[k] for k in SERVING_KEYS if k in serving}
    return doc


def schema():
    with open(SCHEMA_PATH) as f:
        return json.load(f)


def write(path, doc):
    """Writes the document through a temporary file and a rename, as the
    reference publishes its timing artifact."""
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(doc, f)
        f.write("\n")
    os.replace(tmp, path)

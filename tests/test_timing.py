"""Timing documents (nec2c_amd/timing.py): schema version 3 of the reference's
timing artifact (performance.c:150-225), i.e. version 2 plus "hip" as an FFT
backend and an optional serving object.  Checked on the CPU against the
schema file with a small checker that is itself checked on planted faults."""
import json

import pytest

from nec2c_amd import serve, timing, vie
from schema_check import check
import test_serve_host as H

PHASES = [{"name": "fft_build", "seconds": 0.25, "calls": 2}, {"name": "fft_gmres", "seconds": 1.5, "calls": 7},
          {"name": "fft_upload", "seconds": 0.0, "calls": 0}]


def test_document_matches_schema_and_keeps_version_2_layout():
    doc = timing.document("deck.nec", PHASES, 3.0, cpu_seconds=2.0)
    assert check(doc, timing.schema()) == []
    assert [p["name"] for p in doc["phases"]] == ["fft_build", "fft_gmres"]   # no phase without calls
    assert doc["fft_backend"] == {"requested": "hip", "used": "hip"}
    # version 2's keys, in version 2's shapes; only schema_version and the backend value differ
    v2 = dict(doc, schema_version=2, fft_backend={"requested": "builtin", "used": "builtin"})
    assert set(v2) == {"schema", "schema_version", "producer", "input", "dense_backend", "fft_backend",
                       "concurrency", "clock", "completion", "process", "phases"}
    assert json.loads(json.dumps(doc)) == doc


def test_checker_catches_planted_faults():
    sch = timing.schema()
    good = timing.document("d", PHASES, 1.0, cpu_seconds=1.0)
    bad = [dict(good, schema_version=2), dict(good, fft_backend={"requested": "cuda", "used": "hip"}),
           dict(good, extra=1), {k: v for k, v in good.items() if k != "phases"},
           dict(good, phases=[{"name": "x", "seconds": -1.0, "calls": 1}]),
           dict(good, phases=[{"name": "", "seconds": 1.0, "calls": 1}]),
           dict(good, phases=[{"name": "x", "seconds": 1.0, "calls": 0}]),
           dict(good, process=dict(good["process"], peak_rss_kib=1.5))]
    for b in bad:
        assert check(b, sch), b


def test_serving_document_from_the_server_loop(tmp_path):
    built = []
    srv = H.make_server(budget=10 ** 12)
    srv.run([H.request(0, "A", 16, 2, built), H.request(1, "B", 16, 1, built),
             H.request(2, "A", 16, 3, built, arrival=0.0)])
    doc = srv.timing_document()
    assert check(doc, timing.schema()) == []
    s = doc["serving"]
    assert (s["requests"], s["solves"], s["cache_hits"], s["system_builds"]) == (3, 6, 1, 2)
    assert s["p99_ms"] >= s["p50_ms"] >= 0
    out = tmp_path / "timing.json"
    timing.write(str(out), doc)
    assert json.loads(out.read_text()) == json.loads(json.dumps(doc))


def test_library_phases_are_listed():
    ph = vie.phases()
    assert isinstance(ph, list)
    for p in ph:
        assert set(p) == {"name", "seconds", "calls"} and p["calls"] >= 1 and p["name"].startswith("fft_")

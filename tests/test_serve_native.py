"""The library's native serving loop (gpe_serve_* through the C ABI).

The mixed trace of tests/test_serve.py is submitted at once to a server
whose resident-site cap admits only a few requests at a time, so most wait
in the queue and join the batch as others retire.  Every result must equal
the same request run alone in a batch, bit for bit, with the same t and dt;
a request that can never fit the HBM budget is refused at submission."""
import ctypes
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def test_native_server_results_equal_alone():
    from test_serve import trace
    from gpe_amd.batch import Batch
    from gpe_amd.serve import NativeServer
    reqs = trace()
    cap = 2 * max(r.cfg.nx * r.cfg.ny * r.cfg.nz for r in reqs)
    print("[case] native server, %d requests, site cap %d" % (len(reqs), cap), flush=True)
    got, status = {}, {}
    with NativeServer(max_sites=cap) as s:
        tickets = {s.submit(r): r for r in reqs}
        rounds, max_running = 0, 0
        while len(status) < len(reqs):
            running, queued = s.step(3)
            max_running = max(max_running, running)
            for t, st in s.poll():
                status[t] = st
                got[t] = s.result(t)
            rounds += 1
            assert rounds < 10000
        assert running == 0 and queued == 0
    print("[case] %d rounds, at most %d resident" % (rounds, max_running), flush=True)
    assert max_running < len(reqs)  # the cap made requests wait
    assert all(v == 1 for v in status.values()), status
    for t, r in tickets.items():
        with Batch() as b:
            i = r.add_to(b)
            b.run_to_completion(chunk=5)
            want = b.download(i)
            q = b.query(i)
        psi, info = got[t]
        assert info["steps"] == r.steps == q["steps"]
        assert np.array_equal(psi, want), r
        assert info["t"] == q["t"] and info["dt"] == q["dt"]


def test_native_server_refuses_what_never_fits():
    from gpe_amd import _lib
    from gpe_amd.serve import NativeServer
    from test_serve import trace
    r = trace()[0]
    with NativeServer(hbm_budget=1 << 16) as s:  # 64 KiB: below any request's footprint
        with pytest.raises(_lib.GpeError) as e:
            s.submit(r)
        assert e.value.code == -7

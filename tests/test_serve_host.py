"""Server bookkeeping that runs without a device: submission and cancelling
queued requests (no dsmc_server_pump call, so no engine is created)."""
import os
import sys

import pydsmc as D

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dig-appolo_amd"))
import sweep as SW  # noqa: E402


def test_release_of_a_queued_request_cancels_it():
    srv = D.Server(budget_bytes=1 << 40, chunk_steps=5, auto_release=False, batching=4)
    ids = []
    for k in range(3):
        sp = SW.case_spec(k, 0.01)
        m, p = SW.build_case(sp, D.PREC_MIXED, steps=10)
        ids.append(srv.submit(m, p, seed=10 + k, ppc=sp["ppc"], steps=10))
    srv.release(ids[1])
    info = srv.request(ids[1])
    assert info["state"] == D.REQ_FAILED and "cancelled" in info["error"], info
    assert [srv.request(i)["state"] for i in (ids[0], ids[2])] == [D.REQ_QUEUED, D.REQ_QUEUED]
    srv.release(ids[1])  # a second release of the cancelled request changes nothing
    assert srv.request(ids[1])["state"] == D.REQ_FAILED
    srv.close()

"""GPU check of the drop-in device API headers (include/gwf_queue.hpp,
gwf_ring.hpp, glf_queue.hpp, wfqueue_hip_opt.hpp, sfqueue_hip.hpp/.cpp):
the reference harness conventions — single-thread init, FIFO order, exact
capacity and full refusal, empty sentinels, exactly-once delivery under
concurrency — through every API the four harnesses call."""
import pytest

pytestmark = pytest.mark.gpu

APIS = ["gwf_ptr", "gwf_ring", "glf_ptr", "glf_ring", "wfq", "sfq_blocking", "sfq_nb"]


# The rings of the self-check hold 4,096 items; threads stay at or below
# half of that.  Both ring algorithms require threads <= capacity
# (gwf_queue.hpp:320) and, measured, slow down by two orders of magnitude
# right at that bound (4,032 threads: ~4 s against 0.03 s at 3,072).
@pytest.mark.parametrize("api", APIS)
@pytest.mark.parametrize("threads", [256, 2048])
def test_dropin_api(gcq, api, threads):
    print(f"[dropin] {api} threads={threads}", flush=True)
    errs = gcq.dropin_selfcheck(api, threads=threads, ops=16)
    assert all(v == 0 for v in errs.values()), errs

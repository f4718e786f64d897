"""Plane streaming through fmvd_deconvolve_planes_cuda_* with a bounded
look-ahead: the reference's datasource/datasink call order
(fmvd_deconvolve.impl.cpp:216-245,288-294: nStreams gets, then a sink and a
get per plane until a get returns 0, then the remaining sinks) and results
bit-identical to the batched pipeline."""
import numpy as np
import pytest

from gpu_util import make_case

pytestmark = pytest.mark.gpu


def reference_order(D, L):
    """Callback sequence of the reference driver for D >= L planes."""
    seq = [("get", 1)] * L
    z = 0
    for _ in range(L, D):
        seq += [("put", z), ("get", 1)]
        z += 1
    seq += [("put", z), ("get", 0)]
    z += 1
    seq += [("put", k) for k in range(z, D)]
    return seq


@pytest.mark.parametrize("D,L", [(5, 3), (3, 3), (7, 2), (6, 1), (2, 3)])
def test_lookahead_call_order_and_results(fm, D, L):
    V, dh, dw, k = 2, 48, 60, 11
    data, kernels, weights = make_case(D, V, dh, dw, k, k, 16, seed=D * 10 + L)
    planes = [[data[z, v] for v in range(V)] for z in range(D)]
    print(f"\n[lookahead] D={D} L={L}", flush=True)
    with fm.Deconvolver(dh, dw, list(weights), [x.copy() for x in kernels], nstreams=L) as dec:
        batched, best = dec.run(planes, 4, want_estimate=True)
        assert dec.set_lookahead(L) == L
        out, est = dec.run(planes, 4, want_estimate=True)
        calls = dec.calls
    if D >= L:
        assert calls == reference_order(D, L)
    else:  # a source that runs dry during the first L calls ends the stream there
        assert calls == [("get", 1)] * D + [("get", 0)] + [("put", z) for z in range(D)]
    assert np.array_equal(out, batched)
    assert np.array_equal(est, best)
    # at most L planes are ever held between a get and its put
    held = 0
    for kind, ok in calls:
        held += 1 if (kind == "get" and ok) else (-1 if kind == "put" else 0)
        assert 0 <= held <= L

"""The served workload: a Poisson stream of 200 requests/s over three resident meshes (the level-7
icosphere, a 32^3 and a 96^3 jittered Kuhn cube), each request with 1-16 sources and its own
log-normal speed field, solved through eik_batch with one launch in flight at 192 vertices per
block and one sweep per visit (the loop and configuration of bench.py's served leg; the
eik_server path is off).  Every request gets its result; each result of the mix equals the
same request solved alone (a one-solve batch, 1e-5 relative); one 32^3 and one icosphere
request agree with the FP64 oracle (1e-5 relative)."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_err

sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_served_workload_through_batch(E, O):
    nreq = 96
    meshes, fields, reqs, arrivals, _ = bench.serve_workload(E, nreq, 200.0, 4321, vpb=192)
    slot_mesh = [0] * 16 + [1] * 16 + [2] * 8
    b = E.Batch([meshes[k] for k in slot_mesh])
    sample = list(range(0, nreq, 3))
    _say("served stream: %d requests at 200/s over icosphere L7, 32^3, 96^3" % nreq)
    lat, span, launches, kept, _ = bench.batch_serve(b, slot_mesh, reqs, arrivals, fields,
                                                     keep=sample, max_inner=1)
    b.close()
    _say("  %d launches, p50 %.1f ms, p99 %.1f ms" % (launches, 1e3 * np.percentile(lat, 50),
                                                     1e3 * np.percentile(lat, 99)))
    assert (lat > 0).all() and sorted(kept) == sample
    one = [E.Batch([m]) for m in meshes]
    worst = 0.0
    for r in sample:
        k, sd, fi = reqs[r]
        one[k].set_factors(0, fields[k][fi])
        (Ta,), _ = one[k].solve([sd], max_inner=1)
        worst = max(worst, rel_err(kept[r], Ta, 1e-3 * float(Ta.max())))
        assert np.all(kept[r][sd] == 0.0)
    _say("  alone vs mix over %d requests: max rel diff %.3e" % (len(sample), worst))
    assert worst <= 1e-5
    for x in one:
        x.close()

    # the FP64 oracle on the first sampled request of the 32^3 cube and of the icosphere
    r32 = next(r for r in sample if reqs[r][0] == 1)
    k, sd, fi = reqs[r32]
    m = meshes[1]
    xyz, tets = m.vertices().astype(np.float64), m.elements()
    f = fields[1][fi].astype(np.float64)
    with O.exact():
        ref, _ = O.fim3d(xyz, tets, sd, M6=O.tet_metric(tets.shape[0], f))
    err = rel_err(kept[r32], ref, 1e-3 * float(ref.max()))
    _say("  32^3 request %d (%d sources) vs FP64 oracle: %.3e" % (r32, len(sd), err))
    assert err <= 1e-5

    rico = next(r for r in sample if reqs[r][0] == 0)
    k, sd, fi = reqs[rico]
    m = meshes[0]
    xyz, faces = m.vertices().astype(np.float64), m.elements()
    sinv = np.sqrt(fields[0][fi].astype(np.float64))  # factor = 1/speed^2 = speedInv^2
    vf, origin = O.tri_virtual_faces(xyz, faces)
    ref, _ = O.fim2d(xyz, faces, sd, sinv, 0, vfaces=vf, vparam=sinv[origin] if len(vf) else None)
    err = rel_err(kept[rico], ref, 1e-3 * float(ref.max()))
    _say("  icosphere request %d (%d sources) vs FP64 oracle: %.3e" % (rico, len(sd), err))
    assert err <= 1e-5

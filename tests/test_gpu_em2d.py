"""The em2d Python front-end on the GPU: the reference's Python calling
sequences run through the Cython extension and give the same results as the
same set-ups through the C++ objects (pyzpic), and the LWFA window case
through em2d matches the compiled reference's digests bit for bit."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import build_sim

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(msg):
    print(f"[gpu-em2d] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def em2d():
    import em2d
    import pyzpic
    pyzpic.lib()
    return em2d


def test_weibel_through_em2d_matches_pyzpic(em2d):
    import pyzpic
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    uth, ufl = [0.1, 0.1, 0.1], [0.0, 0.0, 0.6]
    el = em2d.Species("electrons", -1.0, [4, 4], udist=em2d.udist.ThermalCorr(uth, ufl))
    po = em2d.Species("positrons", +1.0, [4, 4], udist=em2d.udist.ThermalCorr(uth, [-v for v in ufl]))
    sim = em2d.Simulation([4, 4], [16, 16], [f32(6.4), f32(6.4)], f32(0.07), species=[el, po])
    sim.current.set_filter(em2d.filter.Binomial("x", 1))
    for _ in range(10):
        sim.advance()
    ref = pyzpic.weibel(ntiles=(4, 4), nx=(16, 16), ppc=(4, 4), box=(f32(6.4), f32(6.4)), dt=0.07,
                        filter=("binomial", 0, 1))
    ref.advance(10)
    assert sim.n == 10 and abs(sim.t - ref.get_t()) == 0
    for fc in "xyz":
        assert np.array_equal(sim.emf.E.gather(fc), _gather(ref.field("E")[:, "xyz".index(fc)], ref))
        assert np.array_equal(sim.current.J.gather(fc), _gather(ref.field("J")[:, "xyz".index(fc)], ref))
    for q in ("x", "y", "ux", "uy", "uz"):
        assert np.array_equal(el.gather(q), ref.species[0].gather(q))
    ch = el.get_charge().gather()
    assert ch.shape == (64, 64) and abs(ch.mean() + 1.0) < 1e-5    # electrons at density 1
    pha = el.get_phasespace("ux", [-0.5, 0.5], 64, "uz", [-1.0, 1.0], 32)
    assert pha.shape == (32, 64) and np.isfinite(pha).all()
    e, b = sim.emf.get_energy()
    er = ref.emf_energy()
    assert np.allclose(e, er[0], rtol=0, atol=0) and np.allclose(b, er[1], rtol=0, atol=0)
    ref.close()


def _gather(comp, sim):
    g = sim.geom
    t = comp.reshape(g.ntiles_y, g.ntiles_x, g.tile_vol)
    rows = np.stack([t[:, :, (g.gc_lower + j) * g.ext_x + g.gc_lower:(g.gc_lower + j) * g.ext_x + g.gc_lower + g.nx]
                     for j in range(g.ny)], axis=2)
    return rows.transpose(0, 2, 1, 3).reshape(g.ntiles_y * g.ny, g.ntiles_x * g.nx)


def test_lwfa_window_case_through_em2d_matches_reference(em2d):
    with open(os.path.join(HERE, "golden", "digests_window.json")) as fp:
        entry = json.load(fp)["lwfa_gauss_step"]
    kv = dict(a.split("=", 1) for a in entry["args"])
    f32 = lambda v: float(np.float32(float(v)))  # noqa: E731
    el = em2d.Species("electrons", -1.0, [2, 2], udist=em2d.udist.Cold([0, 0, 0]),
                      density=em2d.density.Step("x", 1.0, f32(2.56)))
    sim = em2d.Simulation([8, 2], [16, 16], [f32(2.56), f32(6.4)], f32(kv["dt"]), species=el, moving_window=True)
    sim.set_deterministic(True)
    p = [f32(v) for v in kv["laser"].split(":")[1].split(",")]
    sim.add(em2d.laser.Gaussian(start=p[0], fwhm=p[1], a0=p[2], omega0=p[3], W0=p[4], focus=p[5], axis=p[6],
                                cos_pol=p[7], sin_pol=p[8]))
    sim.current.set_filter(em2d.filter.Compensated("x", 2))
    # the same run through the C++ objects, state compared with the reference's digests
    import pyzpic
    twin = build_sim(pyzpic, entry["args"], det=True)
    last = max(entry["snap"])
    for step in range(1, last + 1):
        sim.advance()
        twin.advance(1)
    say(f"lwfa via em2d: {el.np_total()} particles after {last} steps")
    E = twin.field("E").ravel()
    assert hashlib.sha256(E.tobytes()).hexdigest() == entry["sha256"][f"s{last}/E"]
    for fc in "xyz":
        assert np.array_equal(sim.emf.E.gather(fc), _gather(twin.field("E")[:, "xyz".index(fc)], twin))
        assert np.array_equal(sim.emf.B.gather(fc), _gather(twin.field("B")[:, "xyz".index(fc)], twin))
    for q in ("x", "y", "ux", "uy", "uz"):
        assert np.array_equal(el.gather(q), twin.species[0].gather(q))
    twin.close()


def test_vec3grid_view_operations(em2d):
    sim = em2d.Simulation([2, 2], [8, 8], [1.6, 1.6], 0.05)
    E = sim.emf.E
    n = E.buffer_size()
    E.set([1.0, 2.0, 3.0])
    assert np.all(E.gather("y") == 2.0)
    E.add(sim.emf.E)
    assert np.all(E.gather("z") == 6.0)
    E.kernel_x(0.25, 0.5, 0.25)                 # constant field: unchanged
    assert np.all(E.gather("x") == 2.0)
    E.x_shift_left(1)                           # periodic x: the field stays constant
    assert np.all(E.gather("x") == 2.0)
    E.zero()
    assert np.all(E.gather("x") == 0.0) and n == 4 * E.buffer_size() // 4

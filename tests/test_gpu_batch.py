"""Batched ensembles: one launch per step over members of different sizes and
parameters, replayed from a captured hipGraph, must give every member exactly
the result of stepping it alone."""
import numpy as np
import pytest

from droplet_amd.abi import DRP_FP32_SHIFTED, DRP_FP64, DRP_WALL_Y, kappa_g_from_sigma, make_config

pytestmark = pytest.mark.gpu


def _members(precision, n=5, seed=4, nu2=0.1):
    rng = np.random.default_rng(seed)
    cfgs, drops = [], []
    for i in range(n):
        N = int(rng.choice([24, 32, 40, 48]))
        cfgs.append(make_config(N, precision=precision, nu1=0.1, nu2=nu2, kappa_g=kappa_g_from_sigma(0.01),
                                wall_axis=DRP_WALL_Y, wall_u=float(rng.uniform(0.005, 0.03)),
                                wall_start=int(rng.integers(1, 9))))
        off = rng.uniform(-2, 2, 3)
        drops.append([tuple(0.5 * (N - 1) + off) + (0.4 * N,)])
    return cfgs, drops


@pytest.mark.parametrize("precision", [DRP_FP64, DRP_FP32_SHIFTED])
@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("nu2", [0.1, 0.2], ids=["nu_const", "nu_phi"])
def test_batch_equals_individual_runs(precision, graph, nu2):
    from droplet_amd.lattice import Batch, Lattice
    cfgs, drops = _members(precision, nu2=nu2)
    b = Batch(cfgs)
    b.use_graph(graph)
    for m, d in zip(b.members, drops):
        m.init_droplets(d)
    b.step(3)
    b.step(17)   # crosses the members' wall start steps, then graph cycles
    b.sync()
    for i, (cfg, d) in enumerate(zip(cfgs, drops)):
        solo = Lattice(cfg)
        solo.init_droplets(d)
        solo.step(20)
        assert b[i].step_count == 20
        a, s = b[i].fields(), solo.fields()
        for k in a:
            np.testing.assert_array_equal(a[k], s[k], err_msg=f"member {i} field {k}")


def test_batch_tile_shapes_and_errors():
    from droplet_amd.lattice import Batch, DropletError
    cfgs, drops = _members(DRP_FP64, n=3)
    b = Batch(cfgs)
    for m, d in zip(b.members, drops):
        m.init_droplets(d)
    b.set_tile_shape(0)
    b.step(7)
    ref = [m.fields()["phi"] for m in b.members]
    b2 = Batch(cfgs)
    for m, d in zip(b2.members, drops):
        m.init_droplets(d)
    b2.step(7)
    for m, r in zip(b2.members, ref):
        np.testing.assert_array_equal(m.fields()["phi"], r)
    b.members[1].step(1)   # members out of step
    with pytest.raises(DropletError):
        b.step(1)
    with pytest.raises(DropletError):
        Batch([make_config(16), make_config(16, precision=DRP_FP32_SHIFTED)])

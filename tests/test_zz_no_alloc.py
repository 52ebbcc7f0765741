"""dem_instance_observables_noalloc, the entry point that allocates no
device memory, against dem_instance_observables on the MI355X.

Two calls in a row on one GPU share its scratch array: a batch of 6 pours,
then a batch of 3 (a smaller input, so the second call reads back fewer
sums than the first left there).  Each result is bitwise the one the
allocating entry point gives on the same state, before and after more
steps; a batch of more instances than the scratch holds is rejected.
"""
import numpy as np
import pytest

from conftest import load_engine
from test_gpu_batch import run_batch, small_batch

pytestmark = pytest.mark.gpu


def test_noalloc_observables_equal_allocating():
    for ninst, seed in ((6, 31), (3, 32)):
        spec, atoms, inst, params = small_batch(ninst=ninst, nmax=1200, seed=seed)
        s = run_batch(spec, atoms, inst, params, 30, graph=True)
        for _ in range(2):
            want = s.observables()
            got = s.observables_noalloc()
            assert got.shape == (ninst, 5)
            assert np.array_equal(got, want), (got, want)
            assert (got[:, 0] == np.bincount(inst, minlength=ninst)).all()
            s.step(20)
        s.close()


def test_noalloc_observables_rejects_too_many_instances():
    """4097 instances (three spheres in a 12 mm box, in instances 0, 1 and
    4096; the rest empty): the allocating entry point serves them, the
    scratch-array one refuses with an error that names its limit."""
    eng = load_engine()
    spec, _, _, _ = small_batch(ninst=1)
    L = 0.012
    spec = dict(spec, box_hi=[L, L, L],
                planes=[(0, 0.0, 1), (0, L, 1), (1, 0.0, 1), (1, L, 1), (2, 0.0, 1), (2, L, 1)])
    r = np.full(3, 1e-3)
    atoms = dict(x=np.array([[0.003, 0.003, 0.003], [0.006, 0.006, 0.006], [0.009, 0.009, 0.009]]),
                 v=np.zeros((3, 3)), omega=np.zeros((3, 3)), radius=r, rmass=4 / 3 * np.pi * r ** 3 * 2500,
                 type=np.ones(3, np.int32), tag=np.arange(1, 4, dtype=np.int32))
    s = eng.Simulation(dict(spec, instances=dict(inst=np.array([0, 1, 4096], np.int32))), atoms)
    s.setup()
    with pytest.raises(RuntimeError, match="scratch limit of 4096"):
        s.observables_noalloc()
    obs = s.observables()
    assert obs.shape == (4097, 5) and obs[:, 0].sum() == 3 and obs[4096, 0] == 1
    s.close()

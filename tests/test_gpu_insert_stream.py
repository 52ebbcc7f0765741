"""fix insert/stream on the MI355X against the oracle (oracle/insert_oracle.c).

The insert_stream tutorial's face, templates, velocity and particle rate
(examples/LIGGGHTS/Tutorials_public/insert_stream/in.insert_stream) in a box
closed below by a wall (the tutorial's shrink-wrapped boundaries would lose
the particles), extrude_length 0.12 (insert_every 6000, 60 bodies per
insertion), gravity, Hertz with tangential history, neigh_modify delay 0:

* at every insertion step the new atoms are the ones the oracle places
  against the device's resident atoms as the step's pre_exchange sees them
  (the previous step's state pushed through the initial half step, written
  here with numpy: the integrator is uncontracted FP64), with the oracle's
  release data, bit for bit;
* after every step an atom awaiting release sits at x_ins + v_n (t - i) dt
  exactly, with v = v_n, no spin, zero force and torque; at its release
  step it takes the insertion velocity;
* graph stepping gives the same run (graphs stay off while atoms await
  release, the integrator un-fused);
* the same in a box periodic in x and y whose x span is the face's (the
  stream's nearby test then runs in the overlap check on each periodic
  image), and with a three-sphere clump template in the distribution,
  placed with random orientations.
"""
import numpy as np
import pytest

from conftest import load_engine
from test_gpu_insert import empty_atoms, materials
from test_insert_stream import tutorial

pytestmark = pytest.mark.gpu

DT = 1e-5


CLUMP_X = np.array([[0.0, 0.0, 0.0], [0.012, 0.0, 0.0], [0.0, 0.01, 0.004]])
CLUMP_R = np.array([0.01, 0.008, 0.006])


def stream_run(case="closed", graph=False):
    eng = load_engine()
    lo, hi, per = [-1.0, -1.0, -0.05], [1.0, 1.0, 2.85], [0, 0, 0]
    if case == "periodic":
        lo, hi, per = [-0.15, 0.1, -0.05], [0.15, 0.55, 2.85], [1, 1, 0]
    spec = dict(box_lo=lo, box_hi=hi, periodic=per, dt=DT, skin=0.01,
                cdf=1.0, neigh_every=1, neigh_delay=0, neigh_check=True, rolling="off",
                materials=materials(Y=5e6, nu=0.45, e=0.9, mu=0.05), planes=[(2, -0.05, 1)],
                gravity=(0.0, 0.0, -9.81), graph=graph)
    ins = tutorial(extrude_length=0.12, nparticles=240)
    s = eng.Simulation(spec, empty_atoms())
    ins_o = ins
    if case == "clumps":
        from oracle import TemplateOracle

        tid = s.add_template_multisphere(15485863, 1, 2500.0, CLUMP_X, CLUMP_R, ntry=20000)
        ins = dict(ins, templates=[dict(type=1, radius=0.015, density=2500.0),
                                   dict(type=1, radius=0.02, density=2500.0)],
                   multisphere=[0, tid], orientation="random")
        ins_o = dict(ins, multisphere_oracle=[None, TemplateOracle(15485863, CLUMP_X, CLUMP_R, ntry=20000)])
    fid = s.add_insert_stream(dict(ins, group_mask=0))
    s.setup()
    return s, fid, ins_o, (lo, hi, per)


def check_pending(s, t):
    """every atom awaiting release at step t rides its release data"""
    st, props = s.state(), s.props()
    rel = s.stream_release_data()[props["tag"] - 1]
    i_step, r_step = rel[:, 3].astype(np.int64), rel[:, 4].astype(np.int64)
    pend = (i_step > 0) & (t < r_step)
    now = (i_step > 0) & (t == r_step)
    te = ((t - i_step[pend]) * DT)[:, None]
    np.testing.assert_array_equal(st["x"][pend], rel[pend, 0:3] + rel[pend, 5:8] * te)
    np.testing.assert_array_equal(st["v"][pend], rel[pend, 5:8])
    assert not st["omega"][pend].any() and not st["f"][pend].any() and not st["torque"][pend].any()
    np.testing.assert_array_equal(st["v"][now], rel[now, 8:11])
    return int(pend.sum())


@pytest.mark.parametrize("case", ["closed", "periodic", "clumps"])
def test_stream_insertion_and_release_match_oracle(case):
    from oracle import InsertStreamOracle

    s, fid, ins, (lo, hi, per) = stream_run(case)
    o = InsertStreamOracle(ins, lo, hi, per, DT)
    dtf = 0.5 * DT
    n_before = 0
    for event in range(4):
        t = int(s.insert_vector(fid)[5])
        assert t == o.next_step()
        if t > 1:
            s.step(t - 1 - s.timestep())
        st, props = s.state(), s.props()
        # the positions pre_exchange sees at step t: FixNVESphere::
        # initial_integrate on the state after step t - 1
        m = props["rmass"][:, None]
        xp = st["x"] + DT * (st["v"] + (dtf / m) * st["f"])
        ref = o.insert(t, xp, props["radius"], props["rmass"])
        s.step(1)
        k = len(ref["radius"])
        st, props = s.state(), s.props()
        print("event %d at step %d: %d resident, %d inserted, device %d" % (event, t, n_before, k, s.n - n_before),
              flush=True)
        assert s.n == n_before + k and k > 0
        new = slice(n_before, n_before + k)
        np.testing.assert_array_equal(props["tag"][new], np.arange(n_before + 1, n_before + k + 1))
        np.testing.assert_array_equal(props["radius"][new], ref["radius"])
        np.testing.assert_array_equal(props["rmass"][new], ref["rmass"])
        rel = s.stream_release_data()[n_before:n_before + k]
        np.testing.assert_array_equal(rel, ref["release"])
        waiting = rel[:, 4] > t
        np.testing.assert_array_equal(st["x"][new][waiting], ref["x"][waiting])
        np.testing.assert_array_equal(s.insert_vector(fid), o.vector())
        assert check_pending(s, t) > 0
        n_before += k
        # mid-way to the next insertion
        s.step(2500)
        check_pending(s, s.timestep())
    assert o.next_step() == 0 and (n_before == 240 or case == "clumps")
    s.step(7000)   # every atom released
    assert check_pending(s, s.timestep()) == 0
    eager = s.state()
    steps = s.timestep()
    s.close()

    g = stream_run(case, graph=True)[0]
    g.step(steps)
    gs = g.state()
    for key in ("x", "v", "omega"):
        np.testing.assert_array_equal(gs[key], eager[key], err_msg=key)
    g.close()

"""Reporters fed from the device (sx_report) on the MI355X.

sx_report copies a frame's fields on the device in stream order and writes
it from a host thread.  Each frame must equal, byte for byte, the frame the
host formatter (pinned to the reference's own output by
tests/test_reporters.py) writes from the state downloaded at the same point
of the step sequence; frames are queued between graph-replayed steps without
waiting, so a frame that read the state late or early would differ.  A
checkpoint read back by sx_load_checkpoint restores the state to the F0.10
rounding of the file and, through the sidecar, the thermostat and step
state.
"""
import numpy as np
import pytest

from oracle.oracle import FIELDS, SCALARS
from tests._states import base_state, to_device

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _lattice_state(rng, n=864, L=10.0776):
    st = base_state(rng, n, L, forces=False, T=1.0)
    g = np.arange(10) * (L / 10) - L / 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)[:n] + rng.normal(size=(n, 3)) * 0.05
    st["rx"], st["ry"], st["rz"] = P[:, 0], P[:, 1], P[:, 2]
    return st


def _host_frame(sx, dev, kind, step, path, rtype, bonds=None):
    fields = dev.state()
    scal = np.array([dev.get_scalar(s.upper()) for s in SCALARS])
    with sx.Reporter(kind, path) as r:
        r.report_host(step, fields, scal, dev.get_box(), rtype=rtype, bonds=bonds)
    return open(path, "rb").read()


def _setup(sx, rng, nbonds=0):
    st = _lattice_state(rng)
    if nbonds:
        st.bbi = np.arange(0, 2 * nbonds, 2)
        st.bbj = st.bbi + 1
        d = np.stack([st[f][st.bbi] - st[f][st.bbj] for f in ("rx", "ry", "rz")], 1)
        d -= np.round(d / st.box[0]) * st.box[0]
        st.bb_len = np.sqrt((d * d).sum(1))
    dev = sx.System(0)
    to_device(dev, st)
    rtype = rng.integers(0, 5, st.n).astype(np.int32)
    dev.set_atom_types(rtype)
    dev.init_nlist(0.3, 2.5)
    dev.initialize_forces()
    dev.forces(sx.LJ, 1.0, 1.0, 2.5)
    return dev, st, rtype


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_device_frames_equal_host_frames(kind, tmp_path):
    import sucrose_amd as sx
    say(f"device-fed reporter kind={kind}: 3 frames between graph-replayed Langevin steps")
    rng = np.random.default_rng(70 + kind)
    spec = [(sx.LJ, 1.0, 1.0, 2.5)]
    a, st, rtype = _setup(sx, rng, nbonds=6 if kind == 4 else 0)
    b, _, _ = _setup(sx, np.random.default_rng(70 + kind), nbonds=6 if kind == 4 else 0)
    bonds = (st.bbi + 1, st.bbj + 1, st.bb_len) if kind == 4 else None
    # a: frames queued between replays without waiting
    rep = a.reporter(kind, tmp_path / "dev.txt")
    if kind == 4:
        rep.set_config_index(3)
    for k in range(3):
        a.run(sx.INT_LGV, spec, 7, 0.004, damping=1.0, first_step=1 + 7 * k)
        rep.report(step=7 * (k + 1))
    rep.close()
    # b: the same steps, the state downloaded after each block, host frames
    host = []
    for k in range(3):
        b.run(sx.INT_LGV, spec, 7, 0.004, damping=1.0, first_step=1 + 7 * k)
        p = tmp_path / f"host{k}.txt"
        if kind == 4:
            with sx.Reporter(kind, p) as r:
                r.set_config_index(3)
                fields = b.state()
                scal = np.array([b.get_scalar(s.upper()) for s in SCALARS])
                r.report_host(7 * (k + 1), fields, scal, b.get_box(), rtype=rtype, bonds=bonds)
            host.append(open(p, "rb").read())
        else:
            host.append(_host_frame(sx, b, kind, 7 * (k + 1), p, rtype))
    if kind == 1:   # one header, then one line per frame
        want = host[0] + b"".join(h.split(b"\n", 1)[1] for h in host[1:])
    else:
        want = b"".join(host)
    got = open(tmp_path / "dev.txt", "rb").read()
    assert got == want
    a.close()
    b.close()


def test_checkpoint_load_restores_state(tmp_path):
    import sucrose_amd as sx
    say("checkpoint: device report, sidecar, sx_load_checkpoint into a fresh context")
    rng = np.random.default_rng(81)
    dev, st, rtype = _setup(sx, rng, nbonds=5)
    dev.set_scalar("ETA1", 0.25)
    dev.set_scalar("PETA2", -1.5)
    dev.run(sx.INT_LGV, [(sx.LJ, 1.0, 1.0, 2.5)], 5, 0.004, damping=1.0, first_step=40)
    p = tmp_path / "chk.txt"
    with dev.reporter(sx.REPORT_CHECKPOINT, p) as r:
        r.set_config_index(9)
        r.report()
    parsed = sx.read_checkpoint(p)
    side = open(str(p) + ".state").read()
    assert "LANGEVIN_STEP 45" in side
    fresh = sx.System(0)
    ci = fresh.load_checkpoint(p)
    assert ci == 9 and fresh.n == st.n
    assert np.array_equal(fresh.get_atom_types(), rtype)
    for f in FIELDS:
        assert np.array_equal(fresh.download(f), parsed["fields"][f.upper()]), f
    assert np.array_equal(fresh.get_box(), parsed["box"])
    for s in ("TARGET_TEMP", "ETA1", "PETA2"):
        assert fresh.get_scalar(s) == dev.get_scalar(s), s
    # the restored system steps on: the same next steps from both
    # (forces recomputed from the rounded positions on both sides)
    q = sx.read_checkpoint(p)
    for f in q["fields"]:
        if f in ("INV_MASS", "IXX_INV", "IYY_INV", "IZZ_INV"):
            continue
        dev.upload(f, q["fields"][f])
    for f, g in (("INV_MASS", "MASS"), ("IXX_INV", "IXX"), ("IYY_INV", "IYY"), ("IZZ_INV", "IZZ")):
        fresh.upload(f, 1.0 / fresh.download(g))
        dev.upload(f, 1.0 / dev.download(g))
    # both continue their Langevin counters (first_step -1): the original
    # from its own, the fresh context from the one the checkpoint restored
    for d in (dev, fresh):
        d.init_nlist(0.3, 2.5)
        d.initialize_forces()
        d.forces(sx.LJ, 1.0, 1.0, 2.5)
        d.run(sx.INT_LGV, [(sx.LJ, 1.0, 1.0, 2.5)], 4, 0.004, damping=1.0, first_step=-1)
    for f in ("RX", "VX", "QW", "LZ"):
        assert np.array_equal(dev.download(f), fresh.download(f)), f
    dev.close()
    fresh.close()

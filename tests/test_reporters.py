"""Reporters (reporters.f90) written by the library's host formatter against
the reference's own output, byte for byte.

The reference, built from its sources (oracle/_ref), writes each frame with
its Fortran runtime; the library formats the same state from host arrays
(sx_report_host).  Where the reference build is absent the comparison runs
against the frames it wrote into tests/golden/reporters/ (make_reporters.py
there).  The checkpoint reader is compared with the reference's
load_checkpoint on the same files.  The device-fed path (sx_report) is
checked against this host formatter in tests/test_gpu_reporters.py.
"""
import os

import numpy as np
import pytest

from oracle.oracle import FIELDS, SCALARS, reference_available
from tests._states import report_state

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reporters")
KINDS = {1: "state_data", 2: "trajectory", 3: "config", 4: "checkpoint"}


def _host_frame(sx, kind, st, rtype, step, path, config_index=0):
    fields = {f: st[f.lower()] for f in FIELDS}
    scal = np.array([st[s] for s in SCALARS])
    bonds = (st.bbi + 1, st.bbj + 1, st.bb_len) if st.nc else None
    with sx.Reporter(kind, path) as r:
        if config_index:
            r.set_config_index(config_index)
        r.report_host(step, fields, scal, st.box, rtype=rtype, bonds=bonds)
    return open(path, "rb").read()


def _cases():
    out = []
    for seed, n, nb in ((1, 20, 0), (2, 37, 12), (3, 5, 4)):
        rng = np.random.default_rng(seed)
        st, rt = report_state(rng, n, nbonds=nb)
        out.append((seed, st, rt))
    return out


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_frames_match_reference_bytes(kind, tmp_path):
    import sucrose_amd as sx
    if reference_available():
        from oracle.oracle import Reference
        ref = Reference(private=True)
    else:
        ref = None
    for seed, st, rt in _cases():
        step = 1000 + seed
        mine = _host_frame(sx, kind, st, rt, step, tmp_path / f"mine_{kind}_{seed}.txt")
        gold = os.path.join(GOLD, f"{KINDS[kind]}_{seed}.txt")
        if ref is not None:
            ref.load(st)
            ref.report(kind, tmp_path / f"ref_{kind}_{seed}.txt", step, rt)
            theirs = open(tmp_path / f"ref_{kind}_{seed}.txt", "rb").read()
            assert os.path.exists(gold) and open(gold, "rb").read() == theirs, "golden file is stale"
        else:
            theirs = open(gold, "rb").read()
        assert mine == theirs, (kind, seed, _first_diff(mine, theirs))


def _first_diff(a, b):
    la, lb = a.decode().splitlines(), b.decode().splitlines()
    for k, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return k, x, y
    return len(la), len(lb)


def test_several_frames_and_header(tmp_path):
    """A reporter writes its header once (state data) and appends frames."""
    import sucrose_amd as sx
    (_, st, rt), (_, st2, rt2) = _cases()[:2]
    p = tmp_path / "sd.txt"
    scal = [np.array([s[k] for k in SCALARS]) for s in (st, st2)]
    with sx.Reporter(sx.REPORT_STATE_DATA, p) as r:
        r.report_host(1, {"rx": st["rx"]}, scal[0], st.box)
        r.report_host(2, {"rx": st2["rx"]}, scal[1], st2.box)
    lines = open(p).read().splitlines()
    assert lines[0].startswith("Step  Potential_Energy") and len(lines) == 3
    assert lines[1].split()[0] == "1" and lines[2].split()[0] == "2"
    p = tmp_path / "traj.txt"
    fields = {f: st[f.lower()] for f in FIELDS}
    with sx.Reporter(sx.REPORT_TRAJECTORY, p) as r:
        for k in range(3):
            r.report_host(k, fields, scal[0], st.box, rtype=rt)
    text = open(p).read()
    assert text.count("ITEM: TIMESTEP") == 3
    assert len(text.splitlines()) == 3 * (9 + st.n)


def test_missing_field_is_an_error(tmp_path):
    import sucrose_amd as sx
    _, st, rt = _cases()[0]
    scal = np.array([st[k] for k in SCALARS])
    with sx.Reporter(sx.REPORT_TRAJECTORY, tmp_path / "t.txt") as r:
        with pytest.raises(sx.SucroseError):
            r.report_host(0, {"rx": st["rx"]}, scal, st.box)


def test_checkpoint_reader_matches_reference(tmp_path):
    """sx_read_checkpoint_host against the reference's load_checkpoint on
    checkpoints the reference wrote (atom and bond lines in shuffled order
    and with a config_index, so that the per-line ids matter), bit for bit."""
    import sucrose_amd as sx
    if not reference_available():
        pytest.skip("reference not built (make -C oracle ref)")
    from oracle.oracle import Reference
    for seed, st, rt in _cases():
        ref = Reference(private=True)
        ref.load(st)
        p = tmp_path / f"chk_{seed}.txt"
        ref.report(4, p, 0, rt)
        lines = open(p).read().split("\n")
        # shuffle the atom lines and the bond lines
        a0 = lines.index("ATOM: ID TYPE MASS CHARGE SHAPEX SHAPEY SHAPEZ ESHAPEX ESHAPEY ESHAPEZ IXX IYY IZZ "
                         "RX RY RZ VX VY VZ LX LY LZ QW QX QY QZ") + 2
        rng = np.random.default_rng(seed)
        atoms = lines[a0:a0 + st.n]
        lines[a0:a0 + st.n] = [atoms[k] for k in rng.permutation(st.n)]
        b0 = lines.index("BOND: ID I J LEN") + 2
        bonds = lines[b0:b0 + st.nc]
        lines[b0:b0 + st.nc] = [bonds[k] for k in rng.permutation(st.nc)]
        lines[-2] = " 17"
        open(p, "w").write("\n".join(lines))
        theirs = Reference(private=True).load_checkpoint(p)
        mine = sx.read_checkpoint(p)
        n, nc, cfg, buf, rtype, bnd, box = theirs
        assert (mine["n"], mine["nc"], mine["config_index"]) == (n, nc, cfg) == (st.n, st.nc, 17)
        assert np.array_equal(mine["box"], box)
        assert np.array_equal(mine["rtype"], rtype)
        for k, f in enumerate(FIELDS):
            assert np.array_equal(mine["fields"][f.upper()], buf[:, k]), f
        for a, b in zip(mine["bonds"], bnd):
            assert np.array_equal(a, b)


def test_checkpoint_round_trip_is_exact_to_ten_decimals(tmp_path):
    """A checkpoint written by the library and read back gives every column
    rounded to 10 decimals (the F0.10 descriptor)."""
    import sucrose_amd as sx
    _, st, rt = _cases()[1]
    p = tmp_path / "c.txt"
    _host_frame(sx, sx.REPORT_CHECKPOINT, st, rt, 0, p, config_index=4)
    back = sx.read_checkpoint(p)
    assert back["config_index"] == 4 and back["n"] == st.n and back["nc"] == st.nc
    assert np.array_equal(back["rtype"], rt)
    for f in ("mass", "charge", "sig_x", "eps_z", "ixx", "rx", "vy", "lz", "qw", "qz"):
        want = np.array([float(f"{x:.10f}") for x in st[f]])
        assert np.array_equal(back["fields"][f.upper()], want), f
    for f in ("inv_mass", "ixx_inv"):
        assert np.all(back["fields"][f.upper()] == -1.0)
    assert np.array_equal(back["bonds"][0], st.bbi + 1)
    assert np.allclose(back["bonds"][2], st.bb_len, atol=5e-11, rtol=0)

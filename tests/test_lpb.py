"""Binary LP container (.nlpb): round trips against the MPS path, an
independent writer of the documented layout, and rejection of truncated,
corrupt and non-canonical files. CPU only (host code of bin/libsov.so,
bin/sovsolve and bin/checker; nothing here touches the GPU)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from nirnaya_amd.nirnaya import NirnayaModel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "bin", "libsov.so")
SOVSOLVE = os.path.join(ROOT, "bin", "sovsolve")
CHECKER = os.path.join(ROOT, "bin", "checker")
CRUDE_MPS = os.path.join(ROOT, "tests", "golden", "MRPL_Crude_Blending_LP.mps")
CRUDE_SOL = os.path.join(ROOT, "tests", "golden", "MRPL_Crude_Blending_LP.sol")


def model():
    return NirnayaModel("t", LIB)


def fnv_words(payload: bytes) -> int:
    """The container's checksum, restated from its documented definition."""
    mask = (1 << 64) - 1
    h = 0xcbf29ce484222325
    n8 = len(payload) // 8 * 8
    for w in np.frombuffer(payload[:n8], dtype="<u8").tolist():
        h = ((h ^ w) * 0x100000001b3) & mask
    for b in payload[n8:]:
        h = ((h ^ b) * 0x100000001b3) & mask
    return h


def write_nlpb(path, m, n, start, index, value, cost, lo, up, rl, ru, sense=1, offset=0.0, checksum=None,
               payload_bytes=None):
    """Independent writer of the documented layout (no names, no types)."""
    arrays = [np.asarray(start, "<i8"), np.asarray(index, "<i4"), np.asarray(value, "<f8"), np.asarray(cost, "<f8"),
              np.asarray(lo, "<f8"), np.asarray(up, "<f8"), np.asarray(rl, "<f8"), np.asarray(ru, "<f8")]
    payload = bytearray()
    at = 128
    for a in arrays:
        pad = (-at) % 64
        payload += bytes(pad)
        at += pad
        payload += a.tobytes()
        at += a.nbytes
    h = fnv_words(bytes(payload)) if checksum is None else checksum
    size = len(payload) if payload_bytes is None else payload_bytes
    header = struct.pack("<8sII4qiidQQ48x", b"NIRLPB01", 128, 0, m, n, len(index), 0, sense, 0, offset, size, h)
    assert len(header) == 128
    with open(path, "wb") as f:
        f.write(header + bytes(payload))


def tiny_arrays():
    # min -x0 - x1, x0 + 2 x1 <= 4, 3 x0 + x1 <= 6, 0 <= x <= 10 (optimum x = (1.6, 1.2), obj -2.8)
    return dict(m=2, n=2, start=[0, 2, 4], index=[0, 1, 0, 1], value=[1.0, 3.0, 2.0, 1.0], cost=[-1.0, -1.0],
                lo=[0.0, 0.0], up=[10.0, 10.0], rl=[-np.inf, -np.inf], ru=[4.0, 6.0])


def test_crude_fixture_round_trip(tmp_path):
    a = model()
    assert a.read_mps(CRUDE_MPS) == 0
    p1 = str(tmp_path / "a.nlpb")
    assert a.write_lpb(p1) == 0
    b = model()
    assert b.read_lpb(p1) == 0
    ma, mb = str(tmp_path / "a.mps"), str(tmp_path / "b.mps")
    assert a.write_mps(ma) == 0 and b.write_mps(mb) == 0
    assert open(ma).read() == open(mb).read()
    # writing again gives the same bytes
    p2 = str(tmp_path / "b.nlpb")
    assert b.write_lpb(p2) == 0
    assert open(p1, "rb").read() == open(p2, "rb").read()
    # and the stored checksum is the documented one
    raw = open(p1, "rb").read()
    assert struct.unpack_from("<Q", raw, 72)[0] == fnv_words(raw[128:])
    assert struct.unpack_from("<Q", raw, 64)[0] == len(raw) - 128


def test_built_model_with_types_and_offset_round_trip(tmp_path):
    a = model()
    a.add_var(1.0, 0.0, 5.0, 0, "x")          # continuous
    a.add_var(-2.0, 0.0, 1.0, 2, "y")         # binary
    a.add_var(0.5, -1e20, 1e20, 1, "z")       # free integer
    a.add_constr([0, 1, 2], [1.0, 1.0, 1.0], -1e20, 3.0, "r0")
    a.add_constr([0, 2], [2.0, -1.0], -1.0, 1e20, "r1")
    a.set_sense(-1)
    a.set_offset(7.25)
    p = str(tmp_path / "m.nlpb")
    assert a.write_lpb(p) == 0
    b = model()
    assert b.read_lpb(p) == 0
    ma, mb = str(tmp_path / "a.mps"), str(tmp_path / "b.mps")
    assert a.write_mps(ma) == 0 and b.write_mps(mb) == 0
    assert open(ma).read() == open(mb).read()


def test_independent_writer_is_read(tmp_path):
    p = str(tmp_path / "t.nlpb")
    write_nlpb(p, **tiny_arrays())
    b = model()
    assert b.read_lpb(p) == 0, b.error_message
    mb = str(tmp_path / "t.mps")
    assert b.write_mps(mb) == 0
    text = open(mb).read()
    assert "RHS" in text and "BOUNDS" in text


@pytest.mark.parametrize("defect,message", [
    ("truncated", "size mismatch"),
    ("flipped", "checksum"),
    ("magic", "magic"),
    ("descending", "ascending"),
    ("start_past_end", "starts"),
    ("zero_value", "zero"),
    ("row_out_of_range", "out of bounds"),
])
def test_corrupt_files_are_rejected(tmp_path, defect, message):
    p = str(tmp_path / "bad.nlpb")
    t = tiny_arrays()
    if defect == "descending":
        t["index"] = [1, 0, 0, 1]
    if defect == "start_past_end":
        t["start"] = [0, 2, 6]
    if defect == "zero_value":
        t["value"] = [1.0, 0.0, 2.0, 1.0]
    if defect == "row_out_of_range":
        t["index"] = [0, 5, 0, 1]
    write_nlpb(p, **t)
    raw = bytearray(open(p, "rb").read())
    if defect == "truncated":
        raw = raw[:-3]
    if defect == "flipped":
        raw[200] ^= 0x10
    if defect == "magic":
        raw[0:8] = b"NIRLPB99"
    open(p, "wb").write(bytes(raw))
    b = model()
    assert b.read_lpb(p) != 0
    assert message in b.error_message


def test_cli_converts_and_checker_reads_binary(tmp_path):
    out = str(tmp_path / "crude.nlpb")
    r = subprocess.run([SOVSOLVE, CRUDE_MPS, "--write-lpb", out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert os.path.getsize(out) > 128
    # the checker certifies the reference's own solution against the binary model
    c = subprocess.run([CHECKER, out, CRUDE_SOL], capture_output=True, text=True, timeout=60)
    c_mps = subprocess.run([CHECKER, CRUDE_MPS, CRUDE_SOL], capture_output=True, text=True, timeout=60)
    assert c.returncode == c_mps.returncode == 0, c.stdout + c.stderr
    assert c.stdout == c_mps.stdout


def test_engine_lp_round_trip_and_min_form(tmp_path):
    from nirnaya_amd.engine import EngineError, gen_mixed_lp, read_lp_binary, write_lp_binary
    lp = gen_mixed_lp(11, 3000, 9000, 6.0, 0.01, 1e3)
    lp.offset = -3.5
    p = str(tmp_path / "g.nlpb")
    write_lp_binary(lp, p)
    back = read_lp_binary(p)
    for name in ("start", "index", "value", "c", "l", "u", "rl", "ru"):
        assert np.array_equal(getattr(lp, name), getattr(back, name), equal_nan=True), name
    assert (back.m, back.n, back.offset) == (lp.m, lp.n, lp.offset)
    # a maximisation comes back in min form: min -c^T x - offset
    a = model()
    a.add_var(2.0, 0.0, 1.0, 0, "x")
    a.add_constr([0], [1.0], -1e20, 1.0, "r")
    a.set_sense(-1)
    a.set_offset(4.0)
    pm = str(tmp_path / "max.nlpb")
    assert a.write_lpb(pm) == 0
    mn = read_lp_binary(pm)
    assert mn.c.tolist() == [-2.0] and mn.offset == -4.0
    # integer models are not engine LPs
    a.add_var(1.0, 0.0, 3.0, 1, "k")
    pi = str(tmp_path / "int.nlpb")
    assert a.write_lpb(pi) == 0
    with pytest.raises(EngineError, match="integer"):
        read_lp_binary(pi)
    # a non-canonical engine LP is refused at write time
    lp.index[1], lp.index[0] = lp.index[0], lp.index[1]
    with pytest.raises(EngineError, match="canonical"):
        write_lp_binary(lp, str(tmp_path / "bad.nlpb"))


def test_full_size_instance_round_trip(tmp_path):
    """The 400,000 x 1,200,000 instance (9.6M nonzeros) through the container."""
    import time
    from nirnaya_amd.engine import gen_mixed_lp, read_lp_binary, write_lp_binary
    lp = gen_mixed_lp(20261, 400_000, 1_200_000, 6.0, 0.01, 1e3)
    p = str(tmp_path / "cfg1.nlpb")
    t0 = time.perf_counter()
    write_lp_binary(lp, p)
    t1 = time.perf_counter()
    back = read_lp_binary(p)
    t2 = time.perf_counter()
    print(f"{os.path.getsize(p) / 1e6:.0f} MB written in {t1 - t0:.2f}s, read in {t2 - t1:.2f}s")
    for name in ("start", "index", "value", "c", "l", "u", "rl", "ru"):
        assert np.array_equal(getattr(lp, name), getattr(back, name)), name


def _presolve_report(path):
    r = subprocess.run([SOVSOLVE, path, "--presolve-report"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    import re
    m = re.search(r"removed (\d+) rows, (\d+) cols, (\d+) tightenings", r.stdout)
    assert m, r.stdout
    print(r.stdout.strip())
    return tuple(int(v) for v in m.groups())


def test_presolve_is_a_no_op_on_the_generated_instances(tmp_path):
    """The facade's presolve (level 2) before the engine: nothing to remove on
    the 2,000 x 5,000 equality LP and the 400,000 x 1,200,000 instance, so the
    engine sees them as generated. The 50,000 x 120,000 family base loses a
    handful of singleton rows and their columns (recorded, not zero)."""
    from nirnaya_amd.engine import gen_equality_lp, gen_mixed_lp, write_lp_binary
    cases = {"cfg0": lambda: gen_equality_lp(20260, 2000, 5000, 5),
             "cfg1": lambda: gen_mixed_lp(20261, 400_000, 1_200_000, 6.0, 0.01, 1e3),
             "cfg4": lambda: gen_equality_lp(20264, 50_000, 120_000, 5)}
    got = {}
    for name, make in cases.items():
        p = str(tmp_path / f"{name}.nlpb")
        write_lp_binary(make(), p)   # no names in the file: the reader supplies C<j> / R<i>
        got[name] = _presolve_report(p)
    assert got["cfg0"] == (0, 0, 0) and got["cfg1"] == (0, 0, 0)
    rows, cols, _ = got["cfg4"]
    assert rows <= 50 and cols <= 50

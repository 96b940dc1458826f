"""libkiwi_io (the host snapshot readers, include/kiwi_io.h) against the reference's own reader
(load_sph_snapshot / detect_snapshot_format, src/snapshot.cpp:17-522, through oracle/_ref/libkiwi_pts.so):
values bit for bit, the rows kept, the first error's message and the warnings on stderr, on valid files
and on every failure the reference reports.  CPU only.
"""
import os
import struct
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import io  # noqa: E402
from oracle import oracle as O  # noqa: E402

pytestmark = pytest.mark.skipif(not os.path.exists(O.PTS_SO), reason="oracle/_ref not built (needs the reference sources)")


def both(path, capfd, fmt="auto", hfact=1.2):
    """(ours, reference) as ("ok", dict, stderr) or ("error", message, stderr)."""
    res = []
    for f in (lambda: io.load_snapshot(path, fmt, hfact), lambda: O.ref_load_snapshot(path, fmt, hfact)):
        try:
            v = ("ok", f())
        except (io.SnapshotError, ValueError) as e:
            v = ("error", str(e))
        res.append(v + (capfd.readouterr().err,))
    return res


def same(path, capfd, **kw):
    a, b = both(path, capfd, **kw)
    assert a[0] == b[0], (a[:2], b[:2])
    if a[0] == "ok":
        assert set(a[1]) == set(b[1])
        for k in a[1]:
            assert np.array_equal(a[1][k], b[1][k]), k
    else:
        assert a[1] == b[1]
    assert a[2] == b[2]
    return a


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return str(p)


def rows9(rng, n):
    v = rng.normal(size=(n, 9))
    v[:, 6:9] = np.abs(v[:, 6:9]) + 0.1
    return v


def plain_text(v, sep=" ", fmt="%.17g", eol="\n"):
    return "".join(sep.join(fmt % x for x in r) + eol for r in v)


def test_plain_ascii_valid_variants(tmp_path, capfd):
    rng = np.random.default_rng(1)
    v = rows9(rng, 300)
    cases = {
        "basic": plain_text(v),
        "crlf_tabs_blank": "\n  \t\n" + plain_text(v, sep="\t ", eol="\r\n") + "\r\n\n",
        "no_final_newline": plain_text(v)[:-1],
        "short_digits_plus_hex": "".join("+" + " ".join("%.6e" % x for x in np.abs(r)) + "\n" for r in v[:50]) +
                                 "0x1.8p1 -0x1p-3 2 3 4 5 6 7 8\n",
        "nul_ends_line": plain_text(v[:3]) + "1 2 3 4 5 6 7 8 9\0 junk\n",
    }
    for name, text in cases.items():
        a = same(write(tmp_path, name, text), capfd)
        assert a[0] == "ok", name
    # every field of a row as the reference reads it
    a = same(write(tmp_path, "b", plain_text(v)), capfd)
    assert np.array_equal(a[1]["id"], np.arange(1, 301)) and np.array_equal(a[1]["rho"], v[:, 8])


@pytest.mark.parametrize("bad", ["1 2 3 4 5 6 7 8\n", "1 2 3 4 5 6 7 8 9 10\n", "1 2 inf 4 5 6 7 8 9\n",
                                 "1 2 3 nan 5 6 7 8 9\n", "1 2 3 4 x 6 7 8 9\n", "1 2 3 4 5 6 7 8 9 #\n"])
def test_plain_ascii_errors(tmp_path, capfd, bad):
    rng = np.random.default_rng(2)
    text = plain_text(rows9(rng, 20)) + "\n" + bad + plain_text(rows9(rng, 5))
    a = same(write(tmp_path, "e", text), capfd)
    assert a[0] == "error" and "line 22" in a[1]


def test_plain_ascii_first_error_across_threads(tmp_path, capfd):
    """60,000 rows parsed by 8 threads: the earliest bad line is reported, not a later thread's."""
    rng = np.random.default_rng(3)
    lines = plain_text(rows9(rng, 60000)).splitlines(True)
    lines[41234] = "1 2 3\n"
    lines[52000] = "1 2 3 4 5 6 7 8 9 10\n"
    lines[17777] = "1 2 3 4 5 6 7 8 nan\n"
    p = write(tmp_path, "t", "".join(lines))
    io.set_threads(8)
    try:
        a = same(p, capfd)
    finally:
        io.set_threads(0)
    assert a[1].startswith("SPH snapshot file (plain_ascii format), line 17778: column 9 is not finite")
    lines[17777] = "1 2 3 4 5 6 7 8 9\n"
    lines[41234] = "1 2 3 4 5 6 7 8 9\n"
    lines[52000] = "1 2 3 4 5 6 7 8 9\n"
    a = same(write(tmp_path, "t2", "".join(lines)), capfd)
    assert a[0] == "ok" and len(a[1]["x"]) == 60000


def test_empty_missing_and_unknown(tmp_path, capfd):
    for text in ("", "\n \n\t\n"):
        a = same(write(tmp_path, "z", text), capfd, fmt="plain_ascii")
        assert a[0] == "error" and "contained no rows" in a[1]
    a = same(str(tmp_path / "missing.txt"), capfd, fmt="plain_ascii")
    assert a[1].startswith("could not open SPH snapshot file")
    a = same(write(tmp_path, "u", "1 2 3 4 5 6 7 8 9\n"), capfd, fmt="bogus")
    assert a[1].startswith("unknown SPH snapshot format 'bogus'")


# ---------------------------------------------------------------- splash_ascii
def splash_text(labels, data, width=24, itype=None, strip=False, preamble=True, eol="\n"):
    """A `splash to ascii` file: a signature line, time/npart/units blocks, the header of left-justified
    labels in fixed fields of `width` characters (plus a bare 'itype' when given), then the rows."""
    out = []
    if preamble:
        out += ["# SPLASH: A visualisation tool for SPH data (c)2004-2024 Daniel Price and contributors",
                "#   time:             time unit", "   1.0000000000000000E+00   3.1557600E+13",
                "#   npart:        gas  dust  sink", f"   {len(data)}   0   0",
                "#   units:     x  y  z", "   3.0856776E+18   3.0856776E+18   3.0856776E+18"]
    head = "# " + "".join(lab.ljust(width - 1)[: width - 1] + " " for lab in labels)
    if itype is not None:
        head += "itype"
    elif strip:
        head = head.rstrip()
    out.append(head)
    for k, r in enumerate(data):
        s = "".join("%24.16E" % x for x in r)
        if itype is not None:
            s += "%10d" % itype[k]
        out.append(s)
    return eol.join(out) + eol


LABELS = ["x [pc]", "y [pc]", "z [pc]", "particle mass [M_{Sun}]", "h [pc]", "density [g/cm^3]", "v_x [km/s]",
          "v_y [km/s]", "v_z [km/s]", "u [erg/g]"]


def test_splash_ascii_variants(tmp_path, capfd):
    rng = np.random.default_rng(4)
    v = np.abs(rng.normal(size=(200, 10))) + 0.01
    it = np.where(rng.random(200) < 0.8, 1, 3)
    cases = {
        "w24_itype": dict(labels=LABELS, data=v, itype=it),
        "w16_itype_crlf": dict(labels=[lab[:15] for lab in LABELS], data=v, width=16, itype=it, eol="\r\n"),
        "w24_no_itype": dict(labels=LABELS, data=v),
        "w24_stripped": dict(labels=LABELS, data=v, strip=True),
        "h_only": dict(labels=[lab for lab in LABELS if not lab.startswith("density")], data=np.delete(v, 5, 1)),
        "rho_only": dict(labels=[lab for lab in LABELS if not lab.startswith("h ")], data=np.delete(v, 4, 1), itype=it),
        "reordered_aliases": dict(labels=["v_z", "mass [M_{Sun}]", "x", "smoothing length [pc]", "y", "z", "vx", "v_y",
                                          "rho"], data=v[:, :9], itype=it),
    }
    for name, kw in cases.items():
        a = same(write(tmp_path, name, splash_text(**kw)), capfd)
        assert a[0] == "ok", (name, a[1])
        assert ("no 'itype' column" in a[2]) == ("itype" not in kw) and ("deriving" in a[2]) == name.endswith("_only")
        assert io.detect_format(str(tmp_path / name)) == O.ref_detect_format(str(tmp_path / name)) == "splash_ascii"
    # hfact reaches the derivations
    same(write(tmp_path, "hf", splash_text(**cases["rho_only"])), capfd, hfact=1.0)
    same(write(tmp_path, "hf2", splash_text(**cases["h_only"])), capfd, hfact=1.3)
    # footer comments and blank lines inside the data block are skipped
    t = splash_text(**cases["w24_itype"]).splitlines(True)
    t.insert(30, "\n")
    t.insert(60, "# a comment\n")
    same(write(tmp_path, "gaps", "".join(t) + "# footer\n"), capfd)


def test_splash_ascii_errors(tmp_path, capfd):
    rng = np.random.default_rng(5)
    v = np.abs(rng.normal(size=(30, 10))) + 0.01
    it = np.ones(30, int)
    bad = {
        "bad_units": dict(labels=[lab.replace("[pc]", "[au]") if lab.startswith("h") else lab for lab in LABELS
                                  if not lab.startswith("density")], data=np.delete(v, 5, 1)),
        "duplicate": dict(labels=["x"] + LABELS[1:9] + ["x"], data=v),
        "missing_vz": dict(labels=LABELS[:8] + ["w", "u"], data=v),
        "no_h_no_rho": dict(labels=[lab for lab in LABELS if not lab.startswith(("h ", "density"))] + ["a", "b"],
                            data=v),
        "all_non_gas": dict(labels=LABELS, data=v, itype=np.full(30, 2)),
        "too_few_cols": dict(labels=LABELS[:7], data=v[:, :7]),
        "bad_width": dict(labels=LABELS, data=v, width=20, itype=it),
    }
    for name, kw in bad.items():
        a = same(write(tmp_path, name, splash_text(**kw)), capfd, fmt="splash_ascii")
        assert a[0] == "error", name
    # the header line alone is enough of a preamble; rows from the first line are not
    a = same(write(tmp_path, "head_only", splash_text(labels=LABELS, data=v, preamble=False)), capfd, fmt="splash_ascii")
    assert a[0] == "ok"
    rows_only = "".join(l for l in splash_text(labels=LABELS, data=v).splitlines(True) if not l.startswith("#"))
    a = same(write(tmp_path, "rows_only", rows_only), capfd, fmt="splash_ascii")
    assert a[0] == "error" and "no header preamble" in a[1]
    rows = splash_text(labels=LABELS, data=v, itype=it).splitlines(True)
    for k, rep in enumerate(("   1.0  2.0\n", rows[12].replace("E+00", "E+00 inf", 1), "nan " * 11 + "\n")):
        t = list(rows)
        t[12 + k] = rep
        a = same(write(tmp_path, f"row{k}", "".join(t)), capfd)
        assert a[0] == "error" and "could not parse a particle row" in a[1]
    for text in ("", "# SPLASH only comments\n# more\n", "# SPLASH\n# a\n 1 2 3\n# b\n"):
        a = same(write(tmp_path, "h", text), capfd, fmt="splash_ascii")
        assert a[0] == "error"


def test_detect_phantom_capture(tmp_path):
    """A file opening with the Phantom capture record (either byte order, 4- or 8-byte defaults) is
    phantom_binary for both detectors; a broken record is not."""
    for ib, rb, end in ((4, 8, "<"), (4, 4, ">"), (8, 8, "<")):
        ic, rc = ("i" if ib == 4 else "q"), ("f" if rb == 4 else "d")
        body = struct.pack(f"{end}{ic}{rc}{ic}{ic}{ic}", 60769, 60878.0, 60878, 1, 690706)
        n = len(body)
        p = write(tmp_path, f"cap{ib}{rb}{end}", struct.pack(f"{end}i", n) + body + struct.pack(f"{end}i", n) + b"\0" * 64)
        assert io.detect_format(p) == O.ref_detect_format(p) == "phantom_binary"
        q = write(tmp_path, f"nocap{ib}{rb}{end}", struct.pack(f"{end}i", n) + body[:-1] + b"\1" + struct.pack(f"{end}i", n))
        assert io.detect_format(q) == O.ref_detect_format(q) == "plain_ascii"


# ---------------------------------------------------------------- phantom_binary
def phantom_dump(blocks, end="<", ib=4, rb=8, ident=b"FT:Phantom test dump", header=None, hdr_nblocks=1,
                 npartoftype=(1, 0), massoftype=(1e-6, 2e-6), cut=None, poke=None):
    """A tagged Phantom dump: capture record, identifier, the eight typed header sections, the block
    count and headers, then every block's arrays.  blocks: list of (n, [(tag, kind, values)]) with
    kind 0..7 = default int, int*1, int*2, int*4, int*8, default real, real*4, real*8."""
    ic, rc = ("i" if ib == 4 else "q"), ("f" if rb == 4 else "d")
    fmt = {0: ic, 1: "b", 2: "h", 3: "i", 4: "q", 5: rc, 6: "f", 7: "d"}

    def rec(payload):
        return struct.pack(end + "i", len(payload)) + payload + struct.pack(end + "i", len(payload))

    out = rec(struct.pack(f"{end}{ic}{rc}{ic}{ic}{ic}", 60769, 60878.0, 60878, 1, 690706))
    out += rec(ident.ljust(100))
    if header is None:
        header = {0: [("nblocks", hdr_nblocks), ("ntypes", len(npartoftype))] + [("npartoftype", v) for v in npartoftype],
                  5: [("massoftype", v) for v in massoftype] + [("hfact", 1.2), ("udist", 3.086e16),
                                                                 ("umass", 1.989e33), ("utime", 3.15e13)]}
    for k in range(8):
        items = header.get(k, [])
        out += rec(struct.pack(end + "i", len(items)))
        if items:
            out += rec(b"".join(t.encode().ljust(16) for t, _ in items))
            out += rec(b"".join(struct.pack(end + fmt[k], v) for _, v in items))
    out += rec(struct.pack(end + "i", len(blocks)))
    for n, arrays in blocks:
        nums = [sum(1 for a in arrays if a[1] == k) for k in range(8)]
        out += rec(struct.pack(end + "q", n) + struct.pack(end + "8i", *nums))
    for n, arrays in blocks:
        for k in range(8):
            for tag, kk, vals in arrays:
                if kk == k:
                    out += rec(tag.encode().ljust(16))
                    out += rec(np.asarray(vals).astype(np.dtype(end + fmt[k])).tobytes())
    if poke is not None:
        out = out[:poke] + bytes([out[poke] ^ 0x40]) + out[poke + 1:]
    return out[:cut] if cut is not None else out


def particle_block(rng, n, vel=True, itype=True, hkind=6, extra=True):
    a = [(c, 7, rng.normal(size=n)) for c in ("x", "y", "z")]
    if vel:
        a += [(c, 7, rng.normal(size=n)) for c in ("vx", "vy", "vz")]
    h = np.abs(rng.normal(size=n)) + 0.01
    h[::17] = 0.0
    h[5::23] = -1.0
    a.append(("h", hkind, h))
    if itype:
        a.append(("itype", 1, np.where(rng.random(n) < 0.85, 1, 2)))
    if extra:
        a.append(("u", 6, rng.random(n)))
    return (n, a)


def test_phantom_binary_variants(tmp_path, capfd):
    rng = np.random.default_rng(6)
    sink = (2, [("x", 7, [0.0, 1.0]), ("m", 7, [1.0, 2.0]), ("h", 7, [0.1, 0.1])])
    cases = {
        "le_i4_r8": dict(blocks=[particle_block(rng, 500), sink]),
        "be_i8_r4": dict(blocks=[particle_block(rng, 300, itype=False), (0, [])], end=">", ib=8, rb=4,
                         npartoftype=(300,), massoftype=(3e-7,)),
        "two_tasks": dict(blocks=[particle_block(rng, 200), sink, particle_block(rng, 150), (0, [])], hdr_nblocks=2),
        "h_real8_no_extra": dict(blocks=[particle_block(rng, 100, hkind=7, extra=False), sink]),
    }
    for name, kw in cases.items():
        p = write(tmp_path, name, phantom_dump(**kw))
        assert io.detect_format(p) == O.ref_detect_format(p) == "phantom_binary"
        a = same(p, capfd)
        assert a[0] == "ok" and "kept" in a[2], name
    units = (3.0857e18, 1.9891e33, 1.0e5)
    p = write(tmp_path, "u", phantom_dump(**cases["le_i4_r8"]))
    a, b = io.load_snapshot(p, units=units), O.ref_load_snapshot(p, units=units)
    assert all(np.array_equal(a[k], b[k]) for k in a)
    capfd.readouterr()
    # a small dump: no velocities; refused unless allowed, then zero with a warning
    small = write(tmp_path, "small", phantom_dump(blocks=[particle_block(rng, 80, vel=False), sink],
                                                  ident=b"ST:Phantom small"))
    a = same(small, capfd)
    assert a[0] == "error" and "allow_missing_velocity" in a[1]
    x = io.load_snapshot(small, allow_missing_velocity=True)
    ours = capfd.readouterr().err
    y = O.ref_load_snapshot(small, allow_missing_velocity=True)
    ref = capfd.readouterr().err
    assert all(np.array_equal(x[k], y[k]) for k in x) and ours == ref and "WARNING" in ours


def test_phantom_binary_errors(tmp_path, capfd):
    rng = np.random.default_rng(7)
    good = dict(blocks=[particle_block(rng, 60), (0, [])])
    full = phantom_dump(**good)
    bad = {
        "ident": dict(good, ident=b"FX:not tagged"),
        "no_hfact": dict(good, header={5: [("massoftype", 1e-6), ("udist", 1.0), ("umass", 1.0), ("utime", 1.0)]}),
        "hdr_blocks": dict(good, hdr_nblocks=3),
        "no_itype_several": dict(blocks=[particle_block(rng, 40, itype=False), (0, [])], npartoftype=(30, 10)),
        "no_gas": dict(blocks=[(3, [("x", 7, [0, 1, 2]), ("y", 7, [0, 1, 2]), ("z", 7, [0, 1, 2]), ("vx", 7, [0] * 3),
                                    ("vy", 7, [0] * 3), ("vz", 7, [0] * 3), ("h", 7, [0.0, -1.0, 0.0])]), (0, [])]),
        "no_h": dict(blocks=[(2, [("x", 7, [0, 1]), ("y", 7, [0, 1]), ("z", 7, [0, 1])]), (0, [])]),
    }
    for name, kw in bad.items():
        a = same(write(tmp_path, name, phantom_dump(**kw)), capfd, fmt="phantom_binary")
        assert a[0] == "error", name
    for cut in (3, 40, 150, 700, len(full) // 2, len(full) - 3):
        a = same(write(tmp_path, f"cut{cut}", full[:cut]), capfd, fmt="phantom_binary")
        assert a[0] == "error", cut
    for poke in (len(full) - 2, 600, 1000):  # a record marker or a size field flipped
        a = same(write(tmp_path, f"poke{poke}", phantom_dump(**good, poke=poke)), capfd, fmt="phantom_binary")
    a = same(write(tmp_path, "plain_as_phantom", "1 2 3 4 5 6 7 8 9\n"), capfd, fmt="phantom_binary")
    assert a[0] == "error" and "capture pattern" in a[1]

"""Tinker file layout of the frame writer (t9_system_write_frame) against the
reference's own dump files test/ref/tinkernist.arc and .uind (first frame:
header, box line, first atom lines and the last atom line, kept as text in
tests/golden/tinkernist_text.json), and archives of several frames read back
by the .xyz reader.  CPU only."""
import json
import os

import numpy as np

import t9hip

KEY = "a-axis 30.0\nneighbor-list\nvdw-cutoff 12.0\newald\newald-cutoff 7.0\n"


def _tinkernist(golden_dir, tmp_path):
    z = np.load(os.path.join(golden_dir, "tinkernist.npz"))
    t = json.load(open(os.path.join(golden_dir, "tinkernist_text.json")))
    n = len(z["types"])
    xyz = z["arc"][0]
    lines = [f"{n}  {t['title']}"]
    for a in range(n):
        bonds = " ".join(str(b + 1) for b in z["i12"][a] if b >= 0)
        lines.append(f"{a + 1} {t['names'][a]} {xyz[a, 0]:.6f} {xyz[a, 1]:.6f} {xyz[a, 2]:.6f} {z['types'][a]} {bonds}")
    p = tmp_path / "water30.xyz"
    p.write_text("\n".join(lines) + "\n")
    return t9hip.System.from_files(str(p), key=KEY), z, t


def _frame_lines(path, n):
    f = open(path).read().splitlines()
    return f[:4] + [f[1 + n]]


def test_arc_and_uind_layout(golden_dir, tmp_path):
    s, z, t = _tinkernist(golden_dir, tmp_path)
    n = s.n
    s.write_frame(tmp_path / "out.arc")
    assert _frame_lines(tmp_path / "out.arc", n) == t["arc"]
    s.write_frame(tmp_path / "out.uind", z["uind"][0])
    assert _frame_lines(tmp_path / "out.uind", n) == t["uind"]


def test_archive_round_trip(golden_dir, tmp_path):
    s, z, _ = _tinkernist(golden_dir, tmp_path)
    n = s.n
    p = tmp_path / "two.arc"
    s.write_frame(p, z["arc"][0])
    s.write_frame(p, z["arc"][1], append=True)
    text = open(p).read().splitlines()
    assert len(text) == 2 * (n + 2)
    for f in range(2):  # each frame is a valid .xyz file for the reader
        q = tmp_path / f"frame{f}.xyz"
        q.write_text("\n".join(text[f * (n + 2):(f + 1) * (n + 2)]) + "\n")
        r = t9hip.System.from_files(str(q), key=KEY)
        np.testing.assert_allclose(r.arrays()["xyz"], z["arc"][f], atol=5e-7)
        np.testing.assert_array_equal(r.arrays()["type"], z["types"])

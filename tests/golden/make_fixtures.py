#!/usr/bin/env python3
"""Write the committed golden fixtures from the reference tree (run once, in a
container where the reference is mounted; the fixtures are then data in git).

What it writes (numbers and data files only, no reference source text):
  * tests/golden/local_frame.xyz, local_frame2.xyz  — inputs of
    test/localframe.cpp and test/localframe3.cpp (test/file/local_frame/)
  * tests/golden/localframe_ref.json — the reference values of those tests:
    every `ref_*` array/scalar of test/localframe.cpp:24-536 and
    test/localframe3.cpp:29-230, keyed "<TEST_CASE>/<SECTION>/<name>"
  * tests/golden/tinkernist.npz — test/ref/tinkernist.{arc,uind,udir,def,tef}:
    2 frames x 2,684 atoms (coordinates Å, dipoles Debye, fields as dumped),
    plus types and 1-2 connectivity of test/file/tinkernist/water30.xyz
  * tests/golden/tinkernist_text.json — the dump layout of the same files: the
    title and atom names of water30.xyz and, of the first frame of
    tinkernist.arc and tinkernist.uind, the header, box and first two atom
    lines and the last atom line (text as written by Tinker's dump routine)
  * tests/golden/aspc_reference.json — the ASPC predictor coefficients of
    src/cu/upredict.cu:50-65 (16 numbers)
  * tests/golden/nacl_vdw_ref.json — test/nacl.cpp NaCl-1 vdW energies,
    gradients, virials and the ion separations of test/file/nacl/
  * tests/golden/c5h12acnh2.xyz, dmso.xyz, hippo19_subset.prm, hippo_ref.json —
    the HIPPO multipole, polarization and charge-transfer tests
    (test/emhippo.cpp, test/ephippo.cpp, test/chgtrn.cpp, test/rephippo.cpp,
    test/ref/emhippo.{1,2}.txt, ephippo.{1,2}.txt, chgtrn.1.txt,
    rephippo.1.txt)
  * tests/golden/h2o10.xyz, water19_records.prm, hippo_water_ref.json — the
    HIPPO repulsion of ten waters (test/rephippo.cpp EREPEL-2,
    test/ref/rephippo.2.txt)
  * tests/golden/hippo19_disp_subset.prm, hippo_disp_ref.json — the HIPPO
    dispersion tests (test/disp.cpp with its own test/file/disp/hippo19.prm,
    test/ref/disp.{1,2}.txt)
  * tinker-gpu_amd/params/amoeba09.prm — the subset of
    test/file/commit_6fe8e913/amoeba09.prm this path uses: the scale-factor
    header and the atom/vdw/bond/angle/ureybrad/multipole/polarize records of
    the atom types in the test systems and the water/ion boxes.
"""
import json
import os
import re
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))

TYPES = {3, 7, 15, 36, 37, 61, 62, 210, 211, 212, 213}
HEADER_KEYS = {"forcefield", "vdwtype", "radiusrule", "radiustype", "radiussize", "epsilonrule", "dielectric",
               "polarization", "opbendtype", "torsionunit", "angle-trigonal"}


def write_prm_subset():
    src = os.path.join(REF, "test/file/commit_6fe8e913/amoeba09.prm")
    lines = open(src).read().splitlines()
    out = ["# Subset of the AMOEBA-2009 parameter set used by the electrostatics path", ""]
    classes = set()
    for ln in lines:
        t = ln.split()
        if len(t) >= 3 and t[0] == "atom" and int(t[1]) in TYPES:
            classes.add(int(t[2]))
    header = True
    i = 0
    while i < len(lines):
        ln = lines[i]
        t = ln.split()
        if header:
            if t and t[0] == "atom":
                header = False
            else:
                if t and len(t) == 2 and (t[0] in HEADER_KEYS or t[0].endswith(("-scale", "-intra", "-cubic",
                                                                                   "-quartic", "-pentic", "-sextic"))):
                    out.append(ln)
                i += 1
                continue
        if not t:
            i += 1
            continue
        k = t[0]
        keep = False
        if k == "atom" and int(t[1]) in TYPES:
            keep = True
        elif k == "vdw" and int(t[1]) in classes:
            keep = True
        elif k == "bond" and int(t[1]) in classes and int(t[2]) in classes:
            keep = True
        elif k in ("angle", "ureybrad") and all(int(x) in classes for x in t[1:4]):
            keep = True
        elif k == "polarize" and int(t[1]) in TYPES:
            keep = True
        elif k == "multipole" and int(t[1]) in TYPES:
            out.append(ln)
            out.extend(lines[i + 1:i + 5])
            i += 5
            continue
        if keep:
            out.append(ln)
        i += 1
    dst = os.path.join(REPO, "tinker-gpu_amd/params/amoeba09.prm")
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    open(dst, "w").write("\n".join(out) + "\n")
    return dst


def parse_test_arrays(path, strip_comments=False):
    """Extract `ref_*` numeric arrays/scalars, tracking TEST_CASE / SECTION."""
    text = open(path).read()
    if strip_comments:  # commented-out alternatives must not replace the live values
        text = re.sub(r"//[^\n]*", "", text)
    out = {}
    case, section = "", ""
    pos = 0
    tok = re.compile(r'TEST_CASE\("([^"]+)"|SECTION\("([^"]+)"\)|const\s+(?:double|int)\s+(ref_\w+)((?:\[\])?(?:\[\d+\])?)\s*=\s*')
    while True:
        m = tok.search(text, pos)
        if not m:
            break
        if m.group(1):
            case, section = m.group(1), ""
            pos = m.end()
            continue
        if m.group(2):
            section = m.group(2).strip()
            pos = m.end()
            continue
        name = m.group(3)
        j = m.end()
        if text[j] == "{":
            depth, k = 0, j
            while True:
                if text[k] == "{":
                    depth += 1
                elif text[k] == "}":
                    depth -= 1
                    if depth == 0:
                        break
                k += 1
            body = text[j:k + 1]
            nums = [float(x) for x in re.findall(r"-?\d+\.?\d*(?:[eE][-+]?\d+)?", body)]
            rows = body.count("{") - 1
            val = np.array(nums).reshape(rows, -1).tolist() if rows > 0 else nums
            pos = k + 1
        else:
            k = text.index(";", j)
            expr = text[j:k]
            parts = [p.strip() for p in expr.split("+")]
            try:
                val = sum(float(p) for p in parts)
            except ValueError:
                val = None
                for p in parts:
                    key = f"{case}/{section}/{p}"
                    if key in out:
                        val = (val or 0.0) + out[key]
            pos = k + 1
        if val is not None:
            out[f"{case}/{section}/{name}"] = val
    return out


def read_dump(path):
    frames, lines = [], open(path).read().splitlines()
    i = 0
    while i < len(lines):
        n = int(lines[i].split()[0])
        i += 2  # title, box
        xyz = np.array([[float(v) for v in lines[i + a].split()[2:5]] for a in range(n)])
        frames.append(xyz)
        i += n
    return np.array(frames)


def write_tinkernist():
    base = os.path.join(REF, "test/ref/tinkernist")
    d = {k: read_dump(base + "." + k) for k in ("arc", "uind", "udir", "def", "tef")}
    lines = open(os.path.join(REF, "test/file/tinkernist/water30.xyz")).read().splitlines()
    n = int(lines[0].split()[0])
    types = np.zeros(n, dtype=np.int32)
    i12 = -np.ones((n, 4), dtype=np.int32)
    for a, ln in enumerate(lines[2:2 + n]):
        t = ln.split()
        types[a] = int(t[5])
        for j, b in enumerate(t[6:]):
            i12[a, j] = int(b) - 1
    np.savez_compressed(os.path.join(HERE, "tinkernist.npz"), types=types, i12=i12, **d)


def write_tinkernist_text():
    lines = open(os.path.join(REF, "test/file/tinkernist/water30.xyz")).read().splitlines()
    n = int(lines[0].split()[0])
    title = lines[0].split(None, 1)[1].strip()
    names = [ln.split()[1] for ln in lines[2:2 + n]]
    out = {"title": title, "names": names}
    for k in ("arc", "uind"):
        f = open(os.path.join(REF, "test/ref/tinkernist." + k)).read().splitlines()
        out[k] = f[:4] + [f[1 + n]]
    json.dump(out, open(os.path.join(HERE, "tinkernist_text.json"), "w"), indent=0)
    return out[k][:2]


def write_aspc():
    """The reference's ASPC predictor table (src/cu/upredict.cu:50-65) as
    numbers: {"aspc": [16 floats]}, most recent solution first."""
    src = open(os.path.join(REF, "src/cu/upredict.cu")).read()
    i = src.index("aspc[16] = {")
    body = src[i:src.index("};", i)]
    vals = [float(a) / float(b) for a, b in re.findall(r"(-?[0-9]+)\.\s*/\s*([0-9]+)\.", body)]
    assert len(vals) == 16, vals
    json.dump({"aspc": vals}, open(os.path.join(HERE, "aspc_reference.json"), "w"), indent=1)
    return vals


def write_nacl():
    """vdW golden values of test/nacl.cpp (NaCl-1, no long-range correction)
    and the ion separations of test/file/nacl/nacl{1,2,3}.xyz."""
    d = parse_test_arrays(os.path.join(REF, "test/nacl.cpp"))
    out = {}
    for sec, f in (("no switch", "nacl1.xyz"), ("switch, near cut", "nacl2.xyz"), ("switch, near off", "nacl3.xyz")):
        lines = open(os.path.join(REF, "test/file/nacl", f)).read().splitlines()
        r = float(lines[2].split()[2])
        pre = f"NaCl-1/- ehal -- {sec}/"
        out[sec] = dict(r=r, eng=d[pre + "ref_eng"], grad=d[pre + "ref_grad"], vir=d[pre + "ref_v"],
                        count=d[pre + "ref_count"])
    json.dump(out, open(os.path.join(HERE, "nacl_vdw_ref.json"), "w"), indent=1)
    return out


def parse_ref_txt(path):
    """A test/ref/*.txt analysis printout: total (or the one component)
    energy, interaction count when printed, virial [3][3], gradient [n][3]."""
    lines = open(path).read().splitlines()
    out = {"grad": []}
    for i, ln in enumerate(lines):
        t = ln.split()
        if "Total Potential Energy" in ln:
            out["energy"] = float(t[-2])
        elif "Energy Component Breakdown" in ln:
            c = lines[i + 1].split()
            out["energy"] = float(c[-2])
            out["count"] = int(c[-1])
        elif "Internal Virial Tensor" in ln:
            out["virial"] = [[float(v) for v in t[-3:]]] + [[float(v) for v in lines[i + k].split()] for k in (1, 2)]
        elif t and t[0] == "Anlyt":
            out["grad"].append([float(v) for v in t[2:5]])
    return out


def copy_file(src_rel, dst_name):
    src = open(os.path.join(REF, src_rel)).read()
    open(os.path.join(HERE, dst_name), "w").write(src)


def write_prm_records(src_rel, dst_name, keep_keys, title):
    """A test-only parameter subset: header scalars plus the named records."""
    lines = open(os.path.join(REF, src_rel)).read().splitlines()
    out = ["# " + title, ""]
    i = 0
    while i < len(lines):
        t = lines[i].split()
        if t and t[0] in ("forcefield", "bond-cubic", "bond-quartic", "angle-cubic", "angle-quartic",
                          "angle-pentic", "angle-sextic", "ureybrad-cubic", "ureybrad-quartic") and len(t) == 2:
            out.append(lines[i])
        elif t and t[0] in keep_keys:
            out.append(lines[i])
            if t[0] == "multipole":
                out.extend(lines[i + 1:i + 5])
                i += 5
                continue
        i += 1
    open(os.path.join(HERE, dst_name), "w").write("\n".join(out) + "\n")


def write_more():
    """Further reference fixtures: polpair (test/polpair.cpp), NaCl-2/NaCl-3
    multipoles (test/nacl.cpp:207-304), the Urey-Bradley golden of ten waters
    (test/urey.cpp) and the bond golden of trp-cage (test/bond.cpp), and the
    argon-box neighbor sets (test/nblist.cpp, test/nblist.hh)."""
    out = {}
    copy_file("test/file/polpair/nacl.xyz", "polpair_nacl.xyz")
    out["polpair_ewald"] = parse_ref_txt(os.path.join(REF, "test/ref/polpair.1.txt"))
    out["polpair_nonewald"] = parse_ref_txt(os.path.join(REF, "test/ref/polpair.2.txt"))
    copy_file("test/file/nacl/nacl1.xyz", "nacl1.xyz")
    copy_file("test/file/nacl/nacl4.xyz", "nacl4.xyz")
    # the multipole records of nacl4.key (its parameter lines, without keywords)
    kl = open(os.path.join(REF, "test/file/nacl/nacl4.key")).read().splitlines()
    recs = []
    for i, ln in enumerate(kl):
        if ln.split()[:1] == ["multipole"]:
            recs.extend(kl[i:i + 5])
    open(os.path.join(HERE, "nacl4_multipoles.key"), "w").write("\n".join(recs) + "\n")
    d = parse_test_arrays(os.path.join(REF, "test/nacl.cpp"), strip_comments=True)
    for case, sec in (("NaCl-2", "empole -- non-ewald, pbc"), ("NaCl-3", "empole -- pme")):
        pre = f"{case}/{sec}/"
        out[case] = dict(energy=d[pre + "ref_eng"], grad=d[pre + "ref_grad"], virial=d[pre + "ref_v"],
                         count=d[pre + "ref_count"])
    copy_file("test/file/water10/h2o10.xyz", "h2o10.xyz")
    write_prm_records("test/file/commit_6fe8e913/water03.prm", "water03_subset.prm",
                      ("atom", "vdw", "bond", "angle", "ureybrad", "multipole", "polarize"),
                      "Subset of the AMOEBA-WATER-2003 parameters (test/file/commit_6fe8e913/water03.prm)")
    out["urey_h2o10"] = parse_ref_txt(os.path.join(REF, "test/ref/urey.txt"))
    copy_file("test/file/trpcage/trpcage.xyz", "trpcage.xyz")
    write_prm_records("test/file/commit_6fe8e913/amoebapro13.prm", "amoebapro13_bond_subset.prm",
                      ("atom", "bond"),
                      "Atom and bond records of AMOEBA-PRO-2013 (test/file/commit_6fe8e913/amoebapro13.prm)")
    out["bond_trpcage"] = parse_ref_txt(os.path.join(REF, "test/ref/bond.txt"))
    json.dump(out, open(os.path.join(HERE, "more_ref.json"), "w"), indent=0)
    # argon box: five frames of coordinates and the reference's neighbor sets
    arc = open(os.path.join(REF, "test/file/arbox/arbox.arc")).read().splitlines()
    n = int(arc[0].split()[0])
    frames, boxes, i = [], [], 0
    while i < len(arc) and len(frames) < 5:
        boxes.append([float(v) for v in arc[i + 1].split()[:3]])
        frames.append([[float(v) for v in arc[i + 2 + a].split()[2:5]] for a in range(n)])
        i += n + 2
    hh = open(os.path.join(REF, "test/nblist.hh")).read()
    body = hh[hh.index("ans[nframe][216][70] = {"):]
    nums = [int(x) for x in re.findall(r"-?\d+", body[body.index("{"):])]
    ans = np.array(nums[:5 * 216 * 70], dtype=np.int16).reshape(5, 216, 70)
    np.savez_compressed(os.path.join(HERE, "arbox_nblist.npz"), xyz=np.array(frames), box=np.array(boxes), ans=ans)
    head = arc[:n + 2]
    open(os.path.join(HERE, "arbox_frame1.xyz"), "w").write("\n".join(head) + "\n")
    return sorted(out)


def write_hippo():
    """HIPPO charge-penetration multipoles, polarization and charge transfer
    (test/emhippo.cpp, test/ephippo.cpp, test/chgtrn.cpp): the coordinates of
    test/file/c5h12acnh2/c5h12acnh2.xyz and test/file/dmso/dmso.xyz, the
    parameter header and the atom, multipole, chgpen, chgtrn, repulsion,
    dispersion and polarize records of their types from
    test/file/hippo/hippo19.prm, and the
    energy, count, virial and gradient of test/ref/emhippo.{1,2}.txt and
    ephippo.{1,2}.txt (Ewald at 7 Å and non-Ewald at 9 Å), chgtrn.1.txt and
    rephippo.1.txt."""
    types = set()
    for rel, dst in (("test/file/c5h12acnh2/c5h12acnh2.xyz", "c5h12acnh2.xyz"), ("test/file/dmso/dmso.xyz", "dmso.xyz")):
        copy_file(rel, dst)
        xyz = open(os.path.join(REF, rel)).read().splitlines()
        types |= {int(ln.split()[5]) for ln in xyz[1:] if len(ln.split()) > 5 and ln.split()[0].isdigit()}
    lines = open(os.path.join(REF, "test/file/hippo/hippo19.prm")).read().splitlines()
    classes = {int(ln.split()[2]) for ln in lines if ln.split()[:1] == ["atom"] and int(ln.split()[1]) in types}
    out = ["# Subset of the HIPPO-2019 parameter set (test/file/hippo/hippo19.prm): header, and the atom,", "# multipole, chgpen, chgtrn, repulsion, dispersion and polarize records of the c5h12acnh2 and dmso", "# test systems", ""]
    i, header = 0, True
    while i < len(lines):
        t = lines[i].split()
        if t and t[0] == "atom":
            header = False
        if header and len(t) >= 1 and not lines[i].lstrip().startswith("#") and re.match(r"^[a-z][a-z0-9-]*$", t[0]) \
                and len(t) <= 2:
            out.append(lines[i])
        elif t and t[0] == "atom" and int(t[1]) in types:
            out.append(lines[i])
        elif t and t[0] == "multipole" and int(t[1]) in types:
            out.extend(lines[i:i + 5])
            i += 5
            continue
        elif t and t[0] in ("chgpen", "chgtrn", "repulsion", "dispersion") and int(t[1]) in classes:
            out.append(lines[i])
        elif t and t[0] == "polarize" and int(t[1]) in types:
            out.append(lines[i])
        i += 1
    open(os.path.join(HERE, "hippo19_subset.prm"), "w").write("\n".join(out) + "\n")
    ref = {"emhippo_ewald": parse_ref_txt(os.path.join(REF, "test/ref/emhippo.1.txt")),
           "emhippo_nonewald": parse_ref_txt(os.path.join(REF, "test/ref/emhippo.2.txt")),
           "ephippo_ewald": parse_ref_txt(os.path.join(REF, "test/ref/ephippo.1.txt")),
           "ephippo_nonewald": parse_ref_txt(os.path.join(REF, "test/ref/ephippo.2.txt")),
           "chgtrn": parse_ref_txt(os.path.join(REF, "test/ref/chgtrn.1.txt")),
           "rephippo": parse_ref_txt(os.path.join(REF, "test/ref/rephippo.1.txt"))}
    json.dump(ref, open(os.path.join(HERE, "hippo_ref.json"), "w"), indent=0)
    return sorted(ref)


def write_hippo_water():
    """HIPPO repulsion of ten waters (test/rephippo.cpp EREPEL-2): the
    coordinates of test/file/water10/h2o10.xyz, the records of
    test/file/hippo/water19.prm (comment lines dropped), and the energy,
    count, virial and gradient of test/ref/rephippo.2.txt."""
    copy_file("test/file/water10/h2o10.xyz", "h2o10.xyz")
    lines = open(os.path.join(REF, "test/file/hippo/water19.prm")).read().splitlines()
    out = ["# Records of the HIPPO water parameter set (test/file/hippo/water19.prm): keywords and parameter records", ""]
    recs = {"atom", "bond", "angle", "ureybrad", "opbend", "strbnd", "polarize", "chgpen", "chgtrn", "dispersion",
            "repulsion", "vdw", "torsion"}
    i = 0
    while i < len(lines):
        t = lines[i].split()
        if t and t[0] == "multipole":
            out.extend(lines[i:i + 5])
            i += 5
            continue
        if t and (t[0] in recs or (re.match(r"^[a-z][a-z0-9-]*$", t[0]) and (len(t) == 2 or t[0] == "d-equals-p")
                                   and not lines[i][0].isspace())):
            out.append(lines[i])
        i += 1
    open(os.path.join(HERE, "water19_records.prm"), "w").write("\n".join(out) + "\n")
    ref = {"rephippo_water": parse_ref_txt(os.path.join(REF, "test/ref/rephippo.2.txt"))}
    json.dump(ref, open(os.path.join(HERE, "hippo_water_ref.json"), "w"), indent=0)
    return sorted(ref)


def write_hippo_disp():
    """HIPPO dispersion (test/disp.cpp): the parameter header and the atom and
    dispersion records of the c5h12acnh2 types from test/file/disp/hippo19.prm,
    and the energy, count, virial and gradient of test/ref/disp.1.txt (5 Å
    cutoff) and disp.2.txt (Ewald dispersion, 5 Å real space)."""
    xyz = open(os.path.join(REF, "test/file/c5h12acnh2/c5h12acnh2.xyz")).read().splitlines()
    types = {int(ln.split()[5]) for ln in xyz[1:] if len(ln.split()) > 5 and ln.split()[0].isdigit()}
    lines = open(os.path.join(REF, "test/file/disp/hippo19.prm")).read().splitlines()
    classes = {int(ln.split()[2]) for ln in lines if ln.split()[:1] == ["atom"] and int(ln.split()[1]) in types}
    out = ["# Subset of the dispersion test's HIPPO-2019 parameter set (test/file/disp/hippo19.prm): header, and",
           "# the atom and dispersion records of the c5h12acnh2 test system", ""]
    header = True
    for ln in lines:
        t = ln.split()
        if t and t[0] == "atom":
            header = False
        if header and len(t) >= 1 and not ln.lstrip().startswith("#") and re.match(r"^[a-z][a-z0-9-]*$", t[0]) \
                and len(t) <= 2:
            out.append(ln)
        elif t and t[0] == "atom" and int(t[1]) in types:
            out.append(ln)
        elif t and t[0] == "dispersion" and is_int_tok(t[1]) and int(t[1]) in classes:
            out.append(ln)
    open(os.path.join(HERE, "hippo19_disp_subset.prm"), "w").write("\n".join(out) + "\n")
    ref = {"disp_cutoff": parse_ref_txt(os.path.join(REF, "test/ref/disp.1.txt")),
           "disp_ewald": parse_ref_txt(os.path.join(REF, "test/ref/disp.2.txt"))}
    json.dump(ref, open(os.path.join(HERE, "hippo_disp_ref.json"), "w"), indent=0)
    return sorted(ref)


def is_int_tok(t):
    try:
        int(t)
        return True
    except ValueError:
        return False


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not mounted")
    print("prm subset:", write_prm_subset())
    for f in ("local_frame.xyz", "local_frame2.xyz"):
        src = open(os.path.join(REF, "test/file/local_frame", f)).read()
        open(os.path.join(HERE, f), "w").write(src)
    ref = {}
    ref.update(parse_test_arrays(os.path.join(REF, "test/localframe.cpp")))
    ref.update(parse_test_arrays(os.path.join(REF, "test/localframe3.cpp")))
    json.dump(ref, open(os.path.join(HERE, "localframe_ref.json"), "w"), indent=0, sort_keys=True)
    print("golden values:", len(ref))
    write_tinkernist()
    print("tinkernist text:", write_tinkernist_text())
    print("aspc table:", write_aspc())
    print("nacl:", write_nacl())
    print("hippo:", write_hippo())
    print("hippo dispersion:", write_hippo_disp())
    print("hippo water:", write_hippo_water())
    print("more:", write_more())


if __name__ == "__main__":
    main()

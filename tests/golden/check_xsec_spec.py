"""Check the oxygen table specification against the reference source text.

hybrid-final_amd/csrc/xsec_oxygen.h describes each tabulated reaction of the
O2 and Ar/O2 chemistries as a row (threshold, excited-target energy shift,
fit or data set, the 3.5 factor of O2(a) attachment), lists the rate
constants, and lists the (projectile, target) pairs of the null-collision
method with their table columns, summation groups, projectile mass and
target density.  This script reads the same facts out of the reference's
src/PhysicsData.c (Oxygen_CrossSectionSET, ArO2_CrossSectionSET) and
src/cuda_mcc.cu (Set_NullCollisionTime_cuda) by pattern matching, without
compiling or running anything, and compares them entry by entry.

Only runs where the reference tree is present:
    python tests/golden/check_xsec_spec.py [/root/reference]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(HERE, "..", "..", "hybrid-final_amd", "csrc", "xsec_oxygen.h")

# fit enumerator -> the reference's function name
REF_FIT = {
    "F_EL": "O2sigmaE", "F_ROT": "O2sigmaR", "F_SD": "O2sigmaSD", "F_SS": "O2sigmaSS", "F_L1": "O2sigmaL1",
    "F_L2": "O2sigmaL2", "F_L3": "O2sigmaL3", "F_L4": "O2sigmaL4", "F_ION": "O2sigmaI", "F_LE": "O2sigmaLE",
    "F_DA": "O2sigmaDA", "F_PD": "O2sigmaPD", "F_DISSIZ": "O2sigmaDISSIZ", "F_EO": "O2sigmaEO",
    "F_E1D": "O2sigmaE1D", "F_E1S": "O2sigmaE1S", "F_E3P0": "O2sigmaE3P0", "F_E5S0": "O2sigmaE5S0",
    "F_E3S0": "O2sigmaE3S0", "F_IO": "O2sigmaIO", "F_ED": "O2sigmaeD", "F_NED": "O2sigmaneD",
    "F_NEGOD": "O2sigmanegOD", "F_DR": "O2sigmaDR", "F_MN": "O2sigmaMN", "F_MNO": "O2sigmaMNO",
    "F_CE": "O2sigmaCE", "F_CEOO": "O2sigmaCEOO", "F_CEO": "O2sigmaCEO", "F_CEO2O": "O2sigmaCEO2O",
    "F_FRAG": "O2sigmaFRAG", "F_NEE": "O2sigmaneE", "F_NEUE": "O2sigmaneuE", "F_POE": "O2sigmapoE",
    "F_POOE": "O2sigmapooE", "F_O2O2": "O2sigmaO2O2", "F_OO": "O2sigmaOO",
}
TKIND = {"T_BG": "BG", "T_FG": "FG", "T_SP": "SP"}


def spec_rows(hdr, name, n):
    body = re.search(r"constexpr Row %s\[%d\] = \{(.*?)\n\};" % (name, n), hdr, re.S).group(1)
    rows = re.findall(r"\{([0-9.]+), (\d), (K_\w+), (\w+), (\d+)\}", body)
    if len(rows) != n:
        raise AssertionError(f"{name}: {len(rows)} rows parsed, expected {n}")
    return rows


def ref_rows(src, fn, data):
    body = src[src.index("void %s(" % fn):]
    body = body[:body.index("Rate constant")]
    out = []
    for b in re.split(r"Rctn = ", body)[1:]:
        r = int(re.match(r"(\d+)", b).group(1))
        th = re.search(r"Th_e\s*=\s*([^;]*);", b).group(1).replace(" ", "")
        rhs = [x for x in re.findall(r"%s\[i\]\.cx_%d = ([^;]*);" % (data, r), b) if x.strip() != "0.0"][0]
        out.append((r, th, rhs, re.findall(r"Dn = (\d+);", b)))
    return out


def check_rows(src, hdr, fn, data, name, n):
    bad = []
    ref = ref_rows(src, fn, data)
    if [r[0] for r in ref] != list(range(n)):
        return [f"{fn}: reactions are not 0..{n - 1}"]
    for (r, th, rhs, dns), (mth, msh, kind, a, b) in zip(ref, spec_rows(hdr, name, n)):
        shift = "" if msh == "0" else "-eShift" + msh
        ok = abs(float(re.sub(r"-eShift\d", "", th)) - float(mth)) < 1e-12
        ok &= (shift in th) if shift else ("eShift" not in th)
        if kind in ("K_FIT", "K_FIT_X35"):
            arg = "energy" + ("" if msh == "0" else "+eShift" + msh)
            fn_ok = re.search(r"\b%s\(%s\)" % (REF_FIT[a], re.escape(arg)), rhs) is not None
            fn_ok |= msh == "0" and "%s(%s[i].xee)" % (REF_FIT[a], data) in rhs
            ok &= fn_ok and rhs.count("O2sigma") == 1 and (("3.5" in rhs) == (kind == "K_FIT_X35"))
        elif kind == "K_VIB":
            ok &= all("O2sigmaE%d(energy)" % k in rhs for k in (1, 2, 3, 4)) and rhs.count("O2sigma") == 4
        elif kind == "K_DATA":
            ok &= dns == [a] and rhs.count("FUNC_decay") == 1
        else:
            ok &= dns == [a, b] and rhs.count("FUNC_decay") == 2 and "-" in rhs
        if not ok:
            bad.append(f"{fn} reaction {r}: reference Th_e={th}, sigma={rhs}, sets={dns}; "
                       f"spec {mth} shift {msh} {kind} {a} {b}")
    return bad


def check_rates(src, hdr, fn, name, first):
    body = src[src.index("void %s(" % fn):]
    body = body[body.index("Rate constant"):body.index("\n}")]
    ref = [(int(i), float(v)) for i, v in
           re.findall(r"Coll_Flag\[(\d+)\]\.RR = Coll_Flag\[\d+\]\.Flag \* ([0-9.e+-]+);", body)]
    mine = [float(x) for x in re.search(r"%s\[\d+\] = \{(.*?)\};" % name, hdr, re.S).group(1).split(",")]
    if [i for i, _ in ref] != list(range(first, first + len(mine))) or [v for _, v in ref] != mine:
        return [f"{name}: reference rate constants {ref} differ from {mine}"]
    return []


def spec_pairs(hdr, name, n):
    body = re.search(r"constexpr Pair %s\[%d\] = \{(.*?)\n\};" % (name, n), hdr, re.S).group(1)
    out = []
    for proj, tk, ti, cols in re.findall(r"\{(\d), (T_\w+), (\d), \{([^}]*)\}\}", body):
        groups, cur = [], []
        for c in (int(x) for x in cols.split(",")):
            if c < 0:
                groups.append(cur)
                cur = []
                if c == -2:
                    break
            else:
                cur.append(c)
        out.append((int(proj), TKIND[tk], int(ti), groups))
    return out


def ref_pairs(src, branch, data, n):
    b = src[src.index(branch):]
    b = b[:b.index("ColProb[0]")]
    out = []
    for k in range(n):
        groups = [[int(x) for x in re.findall(r"%s\[i\]\.cx_(\d+)" % data, m.group(1))]
                  for m in re.finditer(r"sigma\[%d\] \+?= ([^;]*);" % k, b)]
        mass = re.search(r"Host_SigmaV\[%d\]\.val = max\(Host_SigmaV\[%d\]\.val,"
                         r"sqrt\(2\*1\.602e-19\*engy/SP\[(\d)\]\.mass\)\*sigma\[%d\]\)" % (k, k, k), b)
        tgt = re.search(r"Ttarget\[%d\] = 1\.0-exp\(-1 \* dt \* Host_SigmaV\[%d\]\.val ?\* ?(BG|FG|SP)\[(\d)\]"
                        r"\.InitDens\)" % (k, k), b)
        out.append((int(mass.group(1)), tgt.group(1), int(tgt.group(2)), groups))
    return out


def check_pairs(src, hdr, branch, data, name, n):
    mine, ref = spec_pairs(hdr, name, n), ref_pairs(src, branch, data, n)
    return [f"{name} pair {k}: spec {mine[k]}, reference {ref[k]}" for k in range(n) if mine[k] != ref[k]]


def main(ref_root="/root/reference"):
    with open(os.path.join(ref_root, "src", "PhysicsData.c")) as f:
        phys = f.read()
    with open(os.path.join(ref_root, "src", "cuda_mcc.cu")) as f:
        mcc = f.read()
    with open(HDR) as f:
        hdr = f.read()
    bad = []
    bad += check_rows(phys, hdr, "Oxygen_CrossSectionSET", "O2_Data", "kO2Rows", 58)
    bad += check_rows(phys, hdr, "ArO2_CrossSectionSET", "ArO2_Data", "kArO2Rows", 68)
    bad += check_rates(phys, hdr, "Oxygen_CrossSectionSET", "kO2Rates", 58)
    bad += check_rates(phys, hdr, "ArO2_CrossSectionSET", "kArO2Rates", 68)
    bad += check_pairs(mcc, hdr, "MainGas == OXYGEN", "O2_Data", "kO2Pairs", 20)
    bad += check_pairs(mcc, hdr, "MainGas == ARO2", "ArO2_Data", "kArO2Pairs", 25)
    for b in bad:
        print(b)
    print("spec check:", "OK" if not bad else f"{len(bad)} differences")
    return bad


if __name__ == "__main__":
    sys.exit(1 if main(*sys.argv[1:]) else 0)

"""Check the O2 collision tables against the reference source text.

hybrid-final_amd/csrc/o2_chem.h lists, for the O2 chemistry (gas type 1),
each handler group (species, target flag, the sigma*v maximum that scales
its draw) with its reactions in order (running-sum mode, outcome, energy
change, product species), and the target test of each species (time step,
draws, sigma*v index and target density).  This script reads the same facts
out of the reference's src/cuda_mccO2.cu (O2_Electron, O2_O2_ion, O2_O_ion,
O2_O_negative, O2Collision_Check) by pattern matching, without compiling or
running anything, and compares them entry by entry.

Only runs where the reference tree is present:
    python tests/golden/check_o2_chem.py [/root/reference]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(HERE, "..", "..", "hybrid-final_amd", "csrc", "o2_chem.h")

# handler function -> (species, meaning of the reference's Colltype values)
# Colltype meanings follow each handler's second switch: the electron's
# (src/cuda_mccO2.cu:306-425), O2+ (525-572), O+ (668-715), O- (820-864).
HANDLERS = {
    "O2_Electron": (0, {1: ("O2T_LOSS", -1), 2: ("O2T_ATTACH", 3), 3: ("O2T_ION", None), 4: ("O2T_DEL", -1),
                        5: ("O2T_DETACH", 0)}),
    "O2_O2_ion": (1, {1: ("O2T_SCAT", -1), 2: ("O2T_CXKEEP", -1), 3: ("O2T_CXTO", 2)}),
    "O2_O_ion": (2, {1: ("O2T_SCAT", -1), 2: ("O2T_CXTO", 1), 3: ("O2T_CXKEEP", -1)}),
    "O2_O_negative": (3, {1: ("O2T_SCAT", -1), 2: ("O2T_NEGDET", 0), 3: ("O2T_DEL", -1)}),
}
# text the second switch must hold for the product species assumed above
DEST = {
    "O2_Electron": ["case 2: // 2 : Attachment", "data[TID+3*Gsize].PtNumInCell"],
    "O2_O2_ion": ["case 3: // 3 : Charge exchange o+", "atomicAdd(&data[TID+Gsize].PtNumInCell,1)"],
    "O2_O_ion": ["case 2: // 2 : Charge exchange o2+", "atomicAdd(&data[TID-Gsize].PtNumInCell,1)"],
    "O2_O_negative": ["case 2: // 2 : Detachment", "index = info[0].St_num + ID"],
}


def spec(hdr):
    rows = re.findall(r"X\((\d+), (\d), (O2T_\w+), (O2E_\w+), ([0-9.]+)f?, (-?\d), (-?\d)\)",
                      hdr[hdr.index("#define O2_REACTIONS"):])
    groups = re.findall(r"G\((\d), (\d), (\d+), (\d+), (\d+)\)", hdr[hdr.index("#define O2_GROUPS(G)"):])
    targets = re.findall(r"T\((\d), (\d), (\d+), (\d), (\d)\)", hdr[hdr.index("#define O2_TARGETS(T)"):])
    out = {}
    for sp, flag, sig, first, n in groups:
        rr = []
        for r, sm, ty, em, gain, p0, p1 in rows[int(first):int(first) + int(n)]:
            rr.append((int(r), int(sm), ty, em, float(gain) if em == "O2E_GAIN" else 0.0, int(p0), int(p1)))
        out[(int(sp), int(flag))] = (int(sig), rr)
    tg = {(int(a), int(b)): (int(c), int(d), int(e)) for a, b, c, d, e in targets}
    return out, tg


def _fn_body(src, name):
    a = re.search(r"__device__ void\s+%s\(" % name, src).start()
    b = src.index("__device__", a + 20)
    return src[a:b]


def ref_groups(src):
    out = {}
    for fn, (sp, meaning) in HANDLERS.items():
        body = _fn_body(src, fn)
        first_switch = body[body.index("switch(Flag)"):body.index("switch (Colltype)")]
        cases = re.split(r"\n\s*case (\d+):\{", first_switch)
        for i in range(1, len(cases), 2):
            flag, text = int(cases[i]), cases[i + 1]
            sig = int(re.search(r"sigv\[(\d+)\]\.val", text).group(1))
            branches = re.split(r"\}else if|\}else\{", text)
            rows = []
            for br in branches:
                m = re.search(r"R1\s*<=\s*\(SumSigma\s*(\+?=)\s*O2_CrossSection\((\d+),", br)
                if not m:
                    continue
                r = int(m.group(2))
                ct = re.search(r"Colltype\s*=\s*(\d)", br)
                ct = int(ct.group(1)) if ct else (1 if sp == 0 else 0)
                ty, p0 = meaning[ct]
                p1 = -1
                if ty == "O2T_ION":
                    p0 = int(re.search(r"Iz_isp1\s*=\s*(\d)", br).group(1))
                    p1 = int(re.search(r"Iz_isp2\s*=\s*(\d)", br).group(1))
                em, gain = "O2E_NONE", 0.0
                if re.search(r"engy\s*-=\s*info_CX\[%d\]\.Th_e" % r, br):
                    em = "O2E_LOSS"
                g = re.search(r"engy\s*\+=\s*([0-9.]+)f", br)
                if g:
                    em, gain = "O2E_GAIN", float(g.group(1))
                rows.append((r, 1 if m.group(1) == "+=" else 0, ty, em, gain, p0, p1))
            out[(sp, flag)] = (sig, rows)
    return out


def _offset(m):
    """k of "ID+k*Gsize" (0 for "ID", 1 for "ID+Gsize")"""
    return 0 if not m.group(1) else int(m.group(2) or 1)


def ref_targets(src):
    body = _fn_body(src, "O2Collision_Check")
    body = body[:body.index("i = info[isp].St_num")]
    out = {}
    for sp, text in enumerate(re.split(r"\n\s*case \d: //", body)[1:5]):
        dtm = "dtm" in text
        draws_mccn = "Randn = MCCn" in text
        if dtm != (sp <= 1) or draws_mccn != (sp <= 1):
            out[(sp, -1)] = ("time step or draws", dtm, draws_mccn)
        for flag, m in enumerate(re.finditer(r"Prob\d = .*?exp\(-1\*dtm?\*sigv\[(\d+)\]\.val\s*\*\s*([^)]*)\)", text)):
            sig, dens = int(m.group(1)), m.group(2).replace(" ", "")
            if dens.startswith("BG[ID].BackDen1"):
                kind, idx = 0, 0
            elif dens.startswith("Fluid["):
                mm = re.match(r"Fluid\[ID(\+(?:(\d)\*)?Gsize)?\]\.n", dens)
                kind, idx = 1, _offset(mm)
            else:
                mm = re.match(r"data\[ID(\+(?:(\d)\*)?Gsize)?\]\.den\*info\[(\d)\]\.np2c\*dx\*dy", dens)
                kind, idx = 2, int(mm.group(3))
                if _offset(mm) != idx:
                    out[(sp, flag, "den")] = (_offset(mm), idx)
            out[(sp, flag)] = (sig, kind, idx)
    return out


def check(src, hdr):
    bad = []
    mine, tg = spec(hdr)
    ref = ref_groups(src)
    if sorted(mine) != sorted(ref):
        bad.append(f"groups differ: spec {sorted(mine)}, reference {sorted(ref)}")
    for key in sorted(set(mine) & set(ref)):
        if mine[key] != ref[key]:
            bad.append(f"group {key}: spec {mine[key]}, reference {ref[key]}")
    rt = ref_targets(src)
    if tg != rt:
        bad.append(f"targets: spec {tg}, reference {rt}")
    for fn, needles in DEST.items():
        body = _fn_body(src, fn)
        for n in needles:
            if n not in body:
                bad.append(f"{fn}: expected '{n}' in its outcome switch")
    return bad


def main(ref_root="/root/reference"):
    with open(os.path.join(ref_root, "src", "cuda_mccO2.cu")) as f:
        src = f.read()
    with open(HDR) as f:
        hdr = f.read()
    bad = check(src, hdr)
    for b in bad:
        print(b)
    print("O2 table check:", "OK" if not bad else f"{len(bad)} differences")
    return bad


if __name__ == "__main__":
    sys.exit(1 if main(*sys.argv[1:]) else 0)

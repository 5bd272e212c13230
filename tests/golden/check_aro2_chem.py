"""Check the Ar/O2 collision tables against the reference source text.

hybrid-final_amd/csrc/aro2_chem.h lists, for the Ar/O2 chemistry (gas type
2), each handler group (species, target flag, the sigma*v maximum that
scales its draw, the neutral whose velocity the projectile collides in the
frame of) with its reactions in order (running-sum mode, outcome, energy
change and the reaction whose threshold it subtracts, product species), and
the target test of each species (time step, draws, sigma*v index and target
density).  This script reads the same facts out of the reference's
src/cuda_mccArO2.cu (ArO2_Electron, ArO2_Ar_ion, ArO2_O2_ion, ArO2_O_ion,
ArO2_O_negative, ArO2_Collision_Check) by pattern matching, without
compiling or running anything, and compares them entry by entry.

Only runs where the reference tree is present:
    python tests/golden/check_aro2_chem.py [/root/reference]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(HERE, "..", "..", "hybrid-final_amd", "csrc", "aro2_chem.h")

# handler function -> (species, default Colltype, meaning of each Colltype)
# Colltype meanings follow each handler's second switch: the electron's
# (src/cuda_mccArO2.cu:429-556), Ar+ (625-665), O2+ (770-836), O+ (925-979),
# O- (1082-1123).
HANDLERS = {
    "ArO2_Electron": (0, 1, {1: ("O2T_LOSS", -1), 2: ("O2T_ATTACH", 4), 3: ("O2T_IONSHARE", None),
                             4: ("O2T_DEL", -1), 5: ("O2T_DETACH", 0)}),
    "ArO2_Ar_ion": (1, 0, {1: ("O2T_SCAT", -1), 2: ("O2T_CXKEEP", -1)}),
    "ArO2_O2_ion": (2, 0, {1: ("O2T_SCAT", -1), 2: ("O2T_CXKEEP", -1), 3: ("O2T_CXTO", 3), 4: ("O2T_CXTO", 1)}),
    "ArO2_O_ion": (3, 0, {1: ("O2T_SCAT", -1), 2: ("O2T_CXTO", 2), 3: ("O2T_CXKEEP", -1)}),
    "ArO2_O_negative": (4, 0, {1: ("O2T_SCAT", -1), 2: ("O2T_NEGDET", 0), 3: ("O2T_DEL", -1)}),
}
# text the second switch must hold for the product species assumed above
DEST = {
    "ArO2_Electron": ["case 2: // 2 : Attachment", "data[TID+4*Gsize].PtNumInCell", "case 5: // 5 : // Detachment",
                      "data[TID+4*Gsize].PtNumInCell,-1"],
    "ArO2_O2_ion": ["case 3: // 3 : Charge exchange o+", "atomicAdd(&data[TID+Gsize].PtNumInCell,1)",
                    "case 4: // 4 : Charge exchange AR+", "atomicAdd(&data[TID-Gsize].PtNumInCell,1)"],
    "ArO2_O_ion": ["case 2: // 2 : Charge exchange o2+", "atomicAdd(&data[TID-Gsize].PtNumInCell,1)"],
    "ArO2_O_negative": ["case 2: // 2 : Detachment", "index = info[0].St_num + ID"],
}


def spec(hdr):
    rows = re.findall(r"X\((\d+), (\d), (O2T_\w+), (O2E_\w+), ([0-9.]+)f?, (-?\d), (-?\d), (\d+)\)",
                      hdr[hdr.index("#define ARO2_REACTIONS"):])
    groups = re.findall(r"G\((\d), (\d), (\d+), (\d+), (\d+), (\d)\)", hdr[hdr.index("#define ARO2_GROUPS(G)"):])
    targets = re.findall(r"T\((\d), (\d), (\d+), (\d), (\d)\)", hdr[hdr.index("#define ARO2_TARGETS(T)"):])
    out = {}
    for sp, flag, sig, first, n, vsrc in groups:
        rr = []
        for r, sm, ty, em, gain, p0, p1, th in rows[int(first):int(first) + int(n)]:
            rr.append((int(r), int(sm), ty, em, float(gain) if em == "O2E_GAIN" else 0.0, int(p0), int(p1),
                       int(th) if em == "O2E_LOSS" else -1))
        out[(int(sp), int(flag))] = (int(sig), int(vsrc), rr)
    tg = {(int(a), int(b)): (int(c), int(d), int(e)) for a, b, c, d, e in targets}
    return out, tg, len(rows)


def _fn_body(src, name):
    a = re.search(r"__device__ void\s+%s\(" % name, src).start()
    b = src.index("__device__", a + 20)
    return src[a:b]


def _neutral(body, sp, flag):
    """0: no neutral frame; 1/2: the projectile collides in the frame of a
    Maxwellian neutral at BackVel1/BackVel2"""
    pre = body[:body.index("switch(Flag)")]
    if "vneut" not in pre:
        return 0
    m = re.search(r"if\(Flag == 0\) vneut = BG\[ID\]\.BackVel(\d);\s*else\s+vneut = BG\[ID\]\.BackVel(\d);", pre)
    if m:
        return int(m.group(1)) if flag == 0 else int(m.group(2))
    m = re.search(r"dev_maxwellv\(&vneutx,&vneuty,&vneutz,vsave\[n\],BG\[ID\]\.BackVel(\d)", pre)
    return int(m.group(1))


def ref_groups(src):
    out = {}
    for fn, (sp, default, meaning) in HANDLERS.items():
        body = _fn_body(src, fn)
        first_switch = body[body.index("switch(Flag)"):body.index("switch (Colltype)")]
        cases = re.split(r"\n\s*case (\d+):\{", first_switch)
        for i in range(1, len(cases), 2):
            flag, text = int(cases[i]), cases[i + 1]
            sig = int(re.search(r"sigv\[(\d+)\]\.val", text).group(1))
            branches = re.split(r"\}else if|\}else\{", text)
            rows = []
            for br in branches:
                m = re.search(r"R1\s*<=\s*\(SumSigma\s*(\+?=)\s*ArO2_CrossSection\((\d+),", br)
                if not m:
                    continue
                r = int(m.group(2))
                if not re.search(r"engy > info_CX\[%d\]\.Th_e" % r, br):
                    rows.append(("threshold test", r))
                ct = re.search(r"Colltype\s*=\s*(\d)", br)
                ct = int(ct.group(1)) if ct else default
                ty, p0 = meaning[ct]
                p1 = -1
                if ty == "O2T_IONSHARE":
                    p0 = int(re.search(r"Iz_isp1\s*=\s*(\d)", br).group(1))
                    p1 = int(re.search(r"Iz_isp2\s*=\s*(\d)", br).group(1))
                em, gain, th = "O2E_NONE", 0.0, -1
                lm = re.search(r"engy\s*-=\s*info_CX\[(\d+)\]\.Th_e", br)
                if lm:
                    em, th = "O2E_LOSS", int(lm.group(1))
                g = re.search(r"engy\s*\+=\s*([0-9.]+)f", br)
                if g:
                    em, gain = "O2E_GAIN", float(g.group(1))
                if not re.search(r"MCCR\[(TID|ID)\*TnRct\+?%s\]\+\+" % ("" if r == 0 else str(r)), br):
                    rows.append(("event counter", r))
                rows.append((r, 1 if m.group(1) == "+=" else 0, ty, em, gain, p0, p1, th))
            out[(sp, flag)] = (sig, _neutral(body, sp, flag), rows)
    return out


def _offset(m):
    """k of "ID+k*Gsize" (0 for "ID", 1 for "ID+Gsize")"""
    return 0 if not m.group(1) else int(m.group(2) or 1)


def ref_targets(src):
    body = _fn_body(src, "ArO2_Collision_Check")
    body = body[:body.index("i = info[isp].St_num")]
    out = {}
    for sp, text in enumerate(re.split(r"\n\s*case \d: //", body)[1:6]):
        dtm = "dtm" in text
        draws_mccn = "Randn = MCCn" in text
        if dtm != (sp == 0) or draws_mccn != (sp == 0):
            out[(sp, -1)] = ("time step or draws", dtm, draws_mccn)
        for flag, m in enumerate(re.finditer(r"Prob\d = .*?exp\(-1\*dtm?\*sigv\[(\d+)\]\.val\s*\*\s*([^)]*)\)", text)):
            sig, dens = int(m.group(1)), m.group(2).replace(" ", "")
            if dens.startswith("BG[ID].BackDen"):
                kind, idx = 0, int(dens[len("BG[ID].BackDen")]) - 1
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
    mine, tg, nrows = spec(hdr)
    if nrows != 68:
        bad.append(f"the spec lists {nrows} reactions, not 68")
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
    # the dispatch runs every species' handler (TID < nsp*Gsize)
    disp = _fn_body(src.replace("__global__ void", "__device__ void"), "MCC_ArO2_Basic") \
        if "MCC_ArO2_Basic" in src else ""
    if disp and "TID>=nsp*Gsize" not in disp:
        bad.append("MCC_ArO2_Basic: expected the bound TID>=nsp*Gsize")
    return bad


def main(ref_root="/root/reference"):
    with open(os.path.join(ref_root, "src", "cuda_mccArO2.cu")) as f:
        src = f.read()
    with open(os.path.join(ref_root, "src", "cuda_mcc.cu")) as f:
        disp = f.read()
    with open(HDR) as f:
        hdr = f.read()
    bad = check(src + "\n" + disp, hdr)
    for b in bad:
        print(b)
    print("Ar/O2 table check:", "OK" if not bad else f"{len(bad)} differences")
    return bad


if __name__ == "__main__":
    sys.exit(1 if main(*sys.argv[1:]) else 0)

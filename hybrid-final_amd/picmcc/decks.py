"""Case decks in the reference's JSON format (src/start.c:19-1431).

``ccp_deck`` builds a capacitively coupled discharge between two electrodes
that are the bottom and top node rows of the grid: the top one driven with an
RF sine, the bottom one grounded, Neumann side walls.  Keys the reference does
not read are ignored by it; this build adds three under "SimulationMethod":
``LoaderSeed`` (Park-Miller seed of the initial loading, the reference uses
getpid()), ``PhiloxSeed`` (key of the time-loop random stream) and
``RFCycles`` (run length; the reference hard-codes t < 1 ms).
"""
import copy
import json

AR_COLLISIONS = {
    "0.e+Ar>e+Ar": 1.0, "1.e+Ar>e+Ar*": 1.0, "2.e+Ar>e+Ar*m": 1.0, "3.e+Ar>2e+Ar^": 1.0,
    "4.e+Ar*m>2e+Ar^": 1.0, "5.Ar+Ar^>Ar^+Ar": 1.0, "6.Ar+Ar^>Ar+Ar^": 1.0,
    "7.e+Ar*m>e+Ar": 1.0, "8.Ar*m+Ar>e+Ar^+Ar": 1.0, "9.Ar+Ar*m>Ar+Ar": 1.0,
}

KELVIN_TO_EV = 8.6142e-5


def ccp_deck(ncx=64, ncy=64, dx=5e-4, volts=100.0, freq=13.56e6, pres_torr=0.1,
             gas_temp_k=300.0, wall_temp_k=300.0, density=1e15, ppc=50, te_ev=2.5,
             ti_ev=0.026, steps_per_cycle=400, cycles=10, pcg_tol=1e-10, seed=12345,
             philox_seed=20240601, metastable_density=1e15, capacity_factor=3.0):
    """Argon CCP deck; ppc = macro-particles per cell per species."""
    ngx, ngy = ncx + 1, ncy + 1
    np2c = density * dx * dx * 1.0 / ppc
    nmax = int(capacity_factor * ppc * ncx * ncy)
    deck = {
        "GPU_Device_Number": {"GPUDevice": 0},
        "Geometry": {
            "SystemSpec": {"X_length(m)": ncx * dx, "Y_length(m)": ncy * dx, "Z_length(m)": 1.0,
                           "NumGridx": ngx, "NumGridy": ngy},
            "BoundaryCondition": [
                {"X0": 0, "X1": ncx, "Y0": 0, "Y1": 0, "B.C": 1, "Temp": wall_temp_k},
                {"X0": 0, "X1": ncx, "Y0": ncy, "Y1": ncy, "B.C": 1, "Temp": wall_temp_k},
                {"X0": 0, "X1": 0, "Y0": 0, "Y1": ncy, "B.C": 2, "Temp": wall_temp_k},
                {"X0": ncx, "X1": ncx, "Y0": 0, "Y1": ncy, "B.C": 2, "Temp": wall_temp_k},
            ],
            "ConductorSpec": [
                {"M_ID": 1, "X0": 0, "X1": ncx, "Y0": ncy, "Y1": ncy, "Temp": wall_temp_k,
                 "R(Ohm)": 0.0, "L(H)": 0.0, "C(F)": 1.0},
                {"M_ID": 2, "X0": 0, "X1": ncx, "Y0": 0, "Y1": 0, "Temp": wall_temp_k,
                 "R(Ohm)": 0.0, "L(H)": 0.0, "C(F)": 1.0},
            ],
            "Source": [{"M_ID": 1, "DC(V)": 0, "Power(W)": 0, "AC(V)": volts,
                        "Freq(1/s)": freq, "Phase(deg)": 0.0}],
            "DielectricSpec": [],
        },
        "GasSpecies": {
            "Type(0:Ar,1:O2,2:Ar/O2)": 0,
            "TotalPres(Torr)": pres_torr,
            "Background": {"Argon": {"Ratio": 100, "Temp(eV)": gas_temp_k * KELVIN_TO_EV}},
            "NeutralSpecies": {
                "Loadtype": 0,
                "LoadPosition(m)": {"X0": 0.0, "X1": 1.0, "Y0": 0.0, "Y1": 1.0},
                "Ar*": {"Density": metastable_density, "Temp(eV)": 0.026, "Quenching": 1.0},
            },
            "ChargeSpecies": {
                "Loadtype": 0,
                "LoadPosition(m)": {"X0": 0.0, "X1": 1.0, "Y0": 0.0, "Y1": 1.0},
                "Electron": {"S_ID": 0, "Density": density, "Temp(eV)": te_ev, "np2c": np2c,
                             "Ratio": 1, "Max_np": nmax},
                "Ar+": {"S_ID": 1, "Density": density, "Temp(eV)": ti_ev, "np2c": np2c,
                        "Ratio": 1, "Max_np": nmax},
            },
        },
        "SimulationMethod": {
            "TimeStep_PIC": steps_per_cycle, "TimeStep_Conti": 200,
            "FieldTest": {"Laplace_Test_Flag": 0, "Laplace_Field_Solver": 8, "Laplace_Field_Tec_Save": 0},
            "PCGMarginOfError": pcg_tol, "HistoryMax": 1000000, "HistoryDivide": 4,
            "AverageIter": steps_per_cycle, "DensitySmoothing": 0, "ParticleLimit": 2 * nmax,
            "LoaderSeed": seed, "PhiloxSeed": philox_seed, "RFCycles": cycles,
        },
        "Diagnostics": {"BasicDiagnostic": 0},
        "Collision_On_Off": {"ArgonCase": dict(AR_COLLISIONS)},
    }
    return deck


def reaction_keys(gas):
    """(object name under "Collision_On_Off", reaction keys in order) of gas
    type 0 (Ar), 1 (O2) or 2 (Ar/O2), as the library reads them."""
    from . import lib
    txt = lib().pic_reaction_names(gas)
    if txt is None:
        raise ValueError(f"unknown gas type {gas}")
    obj, *keys = txt.decode().split("\n")
    return obj, keys


OXYGEN_NEUTRALS = {"O2A": 1e19, "O2B": 5e18, "OP": 5e19, "OD": 1e18}


def oxygen_deck(gas=2, o2_ratio=0.2, ion_densities=None, neutral_densities=None, ion_collisions=False, **kw):
    """O2 (gas=1) or Ar/O2 (gas=2) CCP deck on the ccp_deck geometry.

    ``o2_ratio`` is the O2 fraction of the background pressure (gas 2; the
    rest is Ar).  Charged-species densities default to a quasi-neutral mix
    (e 1e15, O- 2e14; positive ions make up the rest) with ppc macro-particles
    per cell for e and the molecular/atomic majority ions and ppc/5 for O+
    and O-.  Every reaction switch is on.  Both gas types run on the GPU as
    staged steps."""
    d = ccp_deck(**kw)
    ppc = kw.get("ppc", 50)
    dx = kw.get("dx", 5e-4)
    ncx, ncy = kw.get("ncx", 64), kw.get("ncy", 64)
    g = d["GasSpecies"]
    g["Type(0:Ar,1:O2,2:Ar/O2)"] = gas
    temp = g["Background"]["Argon"]["Temp(eV)"]
    if gas == 1:
        g["Background"] = {"Oxygen": {"Ratio": 1.0, "Temp(eV)": temp}}
        dens = {"Electron": 1e15, "O2+": 1.1e15, "O+": 1e14, "O-": 2e14}
        neut = dict(OXYGEN_NEUTRALS)
    elif gas == 2:
        g["Background"] = {"Argon": {"Ratio": 1.0 - o2_ratio, "Temp(eV)": temp},
                           "Oxygen": {"Ratio": o2_ratio, "Temp(eV)": temp}}
        dens = {"Electron": 1e15, "Ar+": 8e14, "O2+": 3e14, "O+": 1e14, "O-": 2e14}
        neut = {"Ar*": 1e15, **OXYGEN_NEUTRALS}
    else:
        raise ValueError("oxygen_deck builds gas types 1 and 2")
    dens.update(ion_densities or {})
    neut.update(neutral_densities or {})
    ns = g["NeutralSpecies"]
    ns.pop("Ar*", None)
    for name, n in neut.items():
        ns[name] = {"Density": n, "Temp(eV)": 0.026, "Quenching": 1.0}
    cs = g["ChargeSpecies"]
    te, ti = cs["Electron"]["Temp(eV)"], cs["Ar+"]["Temp(eV)"]
    for name in ("Electron", "Ar+"):
        cs.pop(name)
    cf = kw.get("capacity_factor", 3.0)
    for sid, (name, n) in enumerate(dens.items()):
        per_cell = ppc if name in ("Electron", "Ar+", "O2+") else max(1, ppc // 5)
        np2c = n * dx * dx * 1.0 / per_cell
        nmax = int(cf * per_cell * ncx * ncy)
        cs[name] = {"S_ID": sid, "Density": n, "Temp(eV)": te if name == "Electron" else ti,
                    "np2c": np2c, "Ratio": 1, "Max_np": nmax}
    d["SimulationMethod"]["ParticleLimit"] = max(c["Max_np"] for c in cs.values() if isinstance(c, dict)
                                                 and "Max_np" in c)
    obj, keys = reaction_keys(gas)
    d["Collision_On_Off"] = {obj: {k: 1.0 for k in keys}}
    if ion_collisions:
        # O2 only: let O2+, O+ and O- collide too (the reference's O2 step
        # collides electrons only, src/cuda_mcc.cu:76)
        d["SimulationMethod"]["O2IonCollisions"] = 1
    return d


def config3():
    """512x256 Ar/O2 CCP (O2 20 %), dx=0.2 mm, 150 V, 30 mTorr, 100 /cell for e,
    Ar+ and O2+ and 20 /cell for O+ and O-, dt=T/800, 5 cycles."""
    return oxygen_deck(gas=2, o2_ratio=0.2, ncx=512, ncy=256, dx=2e-4, volts=150.0, pres_torr=0.03,
                       ppc=100, steps_per_cycle=800, cycles=5)


def config1():
    """64x64 Ar CCP, dx=0.5 mm, 100 V, 100 mTorr, 50 /cell/species, dt=T/400, 10 cycles."""
    return ccp_deck(ncx=64, ncy=64, dx=5e-4, volts=100.0, pres_torr=0.1, ppc=50,
                    steps_per_cycle=400, cycles=10)


def config2():
    """256x256 Ar CCP, dx=0.2 mm, 200 V, 50 mTorr, 200 /cell/species, dt=T/800, 10 cycles."""
    return ccp_deck(ncx=256, ncy=256, dx=2e-4, volts=200.0, pres_torr=0.05, ppc=200,
                    steps_per_cycle=800, cycles=10, capacity_factor=2.0)


def small(ncx=16, ncy=12, ppc=20, **kw):
    """A small Ar case for parity tests (rectangular on purpose)."""
    base = dict(ncx=ncx, ncy=ncy, dx=5e-4, volts=100.0, pres_torr=0.1, ppc=ppc,
                steps_per_cycle=400, cycles=1)
    base.update(kw)
    return ccp_deck(**base)


def dumps(deck):
    return json.dumps(deck, indent=1)


def with_changes(deck, path, value):
    d = copy.deepcopy(deck)
    cur = d
    for k in path[:-1]:
        cur = cur[k]
    cur[path[-1]] = value
    return d

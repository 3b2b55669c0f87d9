"""Run configuration: the C ``pz_config`` struct and a reader for the
reference's Fortran namelist files.

The namelist groups and defaults follow namelists.f90:162-189 (groups) and
namelists.f90:530-672 (defaults); the string options are mapped to flags the
way read_namelists does (namelists.f90:380-526).
"""
import ctypes
import re

N_T_BOUNDS = 32  # namelists.f90:23

CNAB2, CNLF, MODCNAB, SBDF2, SBDF3, SBDF4, TVB33 = 1, 2, 3, 4, 5, 6, 7
ARS222, ARS443, BPR353, DP1_242, DP2_242, LZ232, CK232, LZ453, PC2 = range(101, 110)
# DIRK schemes with an assembly stage
(SSP222, LSDIRK222, ARS232, ARS233, SSP322, SSP332, SSP433, ARS343, MARS343, KC343, KC564, KC785,
 DBM453) = range(110, 123)
ASSEMBLY_SCHEMES = (SSP222, LSDIRK222, ARS232, ARS233, SSP322, SSP332, SSP433, ARS343, MARS343, KC343,
                    KC564, KC785, DBM453)

# The DIRK names, in the order initialize() of dirk_schemes.f90:55-230 tests
# them.
DIRK_ORDER = [("ARS222", ARS222), ("SSP222", SSP222), ("LSDIRK222", LSDIRK222), ("ARS232", ARS232),
              ("ARS233", ARS233), ("SSP322", SSP322), ("SSP332", SSP332), ("SSP433", SSP433),
              ("DP1-242", DP1_242), ("DP2-242", DP2_242), ("LZ232", LZ232), ("CK232", CK232),
              ("ARS443", ARS443), ("LZ453", LZ453), ("BPR353", BPR353), ("PC2", PC2),
              ("ARS343", ARS343), ("MARS343", MARS343), ("KC343", KC343), ("KC564", KC564),
              ("KC785", KC785), ("DBM453", DBM453)]
# The multistep names in the order of multistep_schemes.f90:56-104
MULTISTEP_ORDER = [("CNAB2", CNAB2), ("MODCNAB", MODCNAB), ("CNLF", CNLF), ("BDF2AB2", SBDF2),
                   ("SBDF2", SBDF2), ("BDF3AB3", SBDF3), ("SBDF3", SBDF3), ("TVB33", TVB33),
                   ("BDF4AB4", SBDF4), ("SBDF4", SBDF4)]
SCHEMES = dict(MULTISTEP_ORDER + [(n, v) for n, v in DIRK_ORDER if v is not None])


def scheme_id(name):
    """The scheme a time_scheme string selects: the family test of
    select_time_scheme.f90:15-49, then the first name initialize() matches."""
    name = str(name).upper()
    order = DIRK_ORDER if any(n in name for n, _ in DIRK_ORDER) else MULTISTEP_ORDER
    for n, v in order:
        if n in name:
            return v
    raise ValueError(f"time scheme {name!r} was not found")

US, UP, OM, PSI, TEMP, XI, TEMP_HAT, XI_HAT, PSI_HAT, DTEMP = range(10)
FIELDS = {"us": US, "up": UP, "om": OM, "psi": PSI, "temp": TEMP, "xi": XI,
          "temp_hat": TEMP_HAT, "xi_hat": XI_HAT, "psi_hat": PSI_HAT,
          "dtemp": DTEMP}


class PzConfig(ctypes.Structure):
    """Mirror of ``struct pz_config`` in include/pizza_amd.h."""
    _fields_ = [
        ("n_r_max", ctypes.c_int), ("n_cheb_max", ctypes.c_int),
        ("m_max", ctypes.c_int), ("minc", ctypes.c_int),
        ("time_scheme", ctypes.c_int), ("l_cheb_coll", ctypes.c_int),
        ("l_direct_solve", ctypes.c_int), ("l_coriolis_imp", ctypes.c_int),
        ("l_buo_imp", ctypes.c_int), ("l_non_rot", ctypes.c_int),
        ("time_scale_rot", ctypes.c_int), ("l_rerror_fix", ctypes.c_int),
        ("rerror_fac", ctypes.c_double), ("alpha", ctypes.c_double),
        ("dtMax", ctypes.c_double), ("courfac", ctypes.c_double),
        ("dt_fac", ctypes.c_double),
        ("ra", ctypes.c_double), ("ek", ctypes.c_double),
        ("pr", ctypes.c_double), ("raxi", ctypes.c_double),
        ("sc", ctypes.c_double), ("radratio", ctypes.c_double),
        ("g0", ctypes.c_double), ("g1", ctypes.c_double),
        ("g2", ctypes.c_double),
        ("ktopt", ctypes.c_int), ("kbott", ctypes.c_int),
        ("ktopv", ctypes.c_int), ("kbotv", ctypes.c_int),
        ("ktopxi", ctypes.c_int), ("kbotxi", ctypes.c_int),
        ("l_ek_pump", ctypes.c_int), ("l_temp_advz", ctypes.c_int),
        ("tcond_fac", ctypes.c_double), ("xicond_fac", ctypes.c_double),
        ("t_top", ctypes.c_double * (3 * N_T_BOUNDS)),
        ("t_bot", ctypes.c_double * (3 * N_T_BOUNDS)),
        ("xi_top", ctypes.c_double * (3 * N_T_BOUNDS)),
        ("xi_bot", ctypes.c_double * (3 * N_T_BOUNDS)),
        ("fd_stretch", ctypes.c_double), ("fd_ratio", ctypes.c_double),
        ("l_finite_diff", ctypes.c_int), ("l_fd_bc_pattern", ctypes.c_int),
    ]


def default_config(**kw):
    """Reference namelist defaults (namelists.f90:530-672), then overrides."""
    c = PzConfig()
    c.n_r_max, c.n_cheb_max, c.m_max, c.minc = 32, 32, 32, 1
    c.time_scheme = CNAB2
    c.l_cheb_coll = 1
    c.l_direct_solve = 1
    c.l_coriolis_imp = 1
    c.l_buo_imp = 1
    c.l_non_rot = 0
    c.time_scale_rot = 0
    c.l_rerror_fix = 0
    c.rerror_fac = 500.0
    c.alpha = 0.5
    c.dtMax = 1.0e-5
    c.courfac = 1.0e3
    c.dt_fac = 2.0
    c.ra, c.ek, c.pr, c.raxi, c.sc = 1.0e5, 1.0e-3, 1.0, 0.0, 0.0
    c.radratio, c.g0, c.g1, c.g2 = 0.35, 0.0, 1.0, 0.0
    c.ktopt = c.kbott = c.ktopxi = c.kbotxi = 1
    c.ktopv = c.kbotv = 2
    c.l_ek_pump = 0
    c.l_temp_advz = 0
    c.tcond_fac = c.xicond_fac = 1.0
    c.fd_stretch, c.fd_ratio = 0.3, 0.1   # namelists.f90:544-545
    c.l_finite_diff = 0
    for k, v in kw.items():
        set_option(c, k, v)
    return c


def _flag(v):
    if isinstance(v, str):
        return 1 if v.strip(".").lower().startswith("t") else 0
    return 1 if v else 0


def set_option(c, key, value):
    """Sets one namelist entry (case-insensitive key) on a PzConfig."""
    k = key.lower()
    if k == "time_scheme":
        c.time_scheme = scheme_id(value)
    elif k == "radial_scheme":
        # namelists.f90:369-376: anything but 'CHEB' selects finite differences
        c.l_finite_diff = 0 if "CHEB" in str(value).upper() else 1
    elif k == "cheb_method":
        c.l_cheb_coll = 1 if "COLL" in str(value).upper() else 0
    elif k == "matrix_solve":
        c.l_direct_solve = 1 if "DIRECT" in str(value).upper() else 0
    elif k == "corio_term":
        c.l_coriolis_imp = 1 if "IMP" in str(value).upper() else 0
    elif k == "buo_term":
        c.l_buo_imp = 1 if "IMP" in str(value).upper() else 0
    elif k == "time_scale":
        c.time_scale_rot = 1 if "ROT" in str(value).upper() else 0
    elif k in ("l_ek_pump", "l_temp_advz", "l_rerror_fix", "l_non_rot"):
        setattr(c, k, _flag(value))
    elif k in ("t_top", "t_bot", "xi_top", "xi_bot"):
        arr = getattr(c, k)
        vals = value if isinstance(value, (list, tuple)) else [value]
        for i, v in enumerate(vals):
            arr[i] = float(v)
    else:
        names = {f[0].lower(): f[0] for f in PzConfig._fields_}
        if k not in names:
            return False
        name = names[k]
        ftype = dict(PzConfig._fields_)[name]
        setattr(c, name, int(value) if ftype is ctypes.c_int else float(value))
    return True


def _parse_value(tok):
    t = tok.strip()
    if not t:
        return None
    if t[0] in "\"'":
        return t.strip("\"'")
    lo = t.lower()
    if lo in (".true.", ".t.", "t", "true"):
        return True
    if lo in (".false.", ".f.", "f", "false"):
        return False
    try:
        return int(t)
    except ValueError:
        return float(lo.replace("d", "e"))


def read_namelist(path):
    """Parses a Fortran namelist file into {group: {key: value}}."""
    text = open(path).read()
    groups = {}
    for g in re.finditer(r"&(\w+)(.*?)\n\s*/", text, re.S):
        body = re.sub(r"!.*", "", g.group(2))
        entries = {}
        for m in re.finditer(r"(\w+)\s*=\s*((?:\"[^\"]*\"|'[^']*'|[^=])*?)(?=,?\s*\w+\s*=|,?\s*$)",
                             body, re.S):
            raw = m.group(2).strip().rstrip(",")
            # list items are separated by commas or blanks (Fortran list-directed input)
            parts = re.findall(r"\"[^\"]*\"|'[^']*'|[^,\s]+", raw)
            vals = [_parse_value(p) for p in parts]
            entries[m.group(1).lower()] = vals[0] if len(vals) == 1 else vals
        groups[g.group(1).lower()] = entries
    return groups


def config_from_namelist(path):
    """Builds (PzConfig, start-field options, output options) from a file."""
    nml = read_namelist(path)
    c = default_config()
    ignored = []
    for grp in ("grid", "control", "phys_param", "hdif"):
        for k, v in nml.get(grp, {}).items():
            if not set_option(c, k, v):
                ignored.append(k)
    start = {"init_t": 0, "amp_t": 0.0, "init_xi": 0, "amp_xi": 0.0,
             "init_u": 0, "amp_u": 0.0, "l_start_file": False}
    start.update(nml.get("start_field", {}))
    control = nml.get("control", {})
    out = {"n_log_step": 50, "n_time_steps": int(control.get("n_time_steps", 100)),
           "tag": control.get("tag", "test")}
    out.update({k: v for k, v in nml.get("output_control", {}).items()})
    return c, start, out

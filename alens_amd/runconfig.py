"""The reference's run configuration file (RunConfig.yaml) for Context.

TubuleSystem reads RunConfig.yaml into SylinderSystem's runConfig when it starts (SRC/main.cpp:26,
SRC/TubuleSystem.cpp:79); the keys are listed in docs/source/quickstart.md:316-358.  load() returns
the keyword arguments of alens_amd.capi.Context for the keys this engine uses, the same mapping as
the C++ adapter's RunConfig::toEngineConfig (include/alens/RunConfig.hpp):
simBoxLow/High/PBC -> box_lo/box_hi/pbc, monolayer, sylinderColBuf -> col_buf, viscosity,
KBT -> kT (<= 0: no Brownian motion), dt, conResTol -> tol, conMaxIte -> max_ite,
conSolverChoice -> solver, sylinderDiameterColRatio / sylinderLengthColRatio -> the collision ratios,
rngSeed -> seed.
"""
import yaml

REQUIRED = ("simBoxLow", "simBoxHigh", "simBoxPBC", "viscosity", "KBT", "dt", "sylinderColBuf", "conResTol",
            "conMaxIte")


def _num(v, key):
    if isinstance(v, bool):
        raise ValueError(f"RunConfig: key '{key}' is not a number: {v!r}")
    try:
        return float(v)  # YAML 1.1 reads 1e-5 (no dot) as a string
    except (TypeError, ValueError):
        raise ValueError(f"RunConfig: key '{key}' is not a number: {v!r}") from None


def _int(v, key):
    x = _num(v, key)
    if x != int(x):
        raise ValueError(f"RunConfig: key '{key}' is not an integer: {v!r}")
    return int(x)


def read(path):
    """The file's keys as a dict (PyYAML safe loader)."""
    with open(path) as fh:
        d = yaml.safe_load(fh)
    if not isinstance(d, dict):
        raise ValueError(f"{path}: not a mapping of keys")
    return d


def load(path, **engine):
    """Context keyword arguments from a RunConfig.yaml file; `engine` adds the settings that are not
    reference keys (precision, warm_start, device, cell_size)."""
    d = read(path)
    for k in REQUIRED:
        if k not in d:
            raise ValueError(f"RunConfig: required key '{k}' is missing")

    def vec(key, conv):
        v = d[key]
        if not isinstance(v, (list, tuple)) or len(v) != 3:
            raise ValueError(f"RunConfig: key '{key}' needs 3 items")
        return tuple(conv(x, key) for x in v)

    def boolean(x, key):
        if not isinstance(x, bool):
            raise ValueError(f"RunConfig: key '{key}' is not a boolean: {x!r}")
        return int(x)

    cfg = dict(box_lo=vec("simBoxLow", _num), box_hi=vec("simBoxHigh", _num), pbc=vec("simBoxPBC", boolean),
               monolayer=boolean(d.get("monolayer", False), "monolayer"), col_buf=_num(d["sylinderColBuf"], "sylinderColBuf"),
               viscosity=_num(d["viscosity"], "viscosity"), kT=_num(d["KBT"], "KBT"), dt=_num(d["dt"], "dt"),
               tol=_num(d["conResTol"], "conResTol"), max_ite=_int(d["conMaxIte"], "conMaxIte"),
               solver=_int(d.get("conSolverChoice", 0), "conSolverChoice"),
               diam_col_ratio=_num(d.get("sylinderDiameterColRatio", 1.0), "sylinderDiameterColRatio"),
               len_col_ratio=_num(d.get("sylinderLengthColRatio", 1.0), "sylinderLengthColRatio"),
               seed=_int(d.get("rngSeed", 0), "rngSeed"))
    cfg.update(engine)
    return cfg

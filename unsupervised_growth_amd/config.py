"""The reference's configuration files (libconfig syntax).

growth_parameters.cfg: ConfigurationNetworkGrowth (c++/src/ConfigurationNetworkGrowth.cpp:
reader 7-131, writer 210-296) -- groups synapticParameters, maturationParameters,
noiseParameters, connectionParameters.  network_parameters.cfg:
ConfigurationNetworkTopology (c++/src/ConfigurationNetworkTopology.cpp) -- group
spatialParameters.  The writers reproduce the reference's text (fixed notation
with 6 decimals from A_P on, as std::fixed / setprecision(6) leave the stream);
the reader takes the subset of libconfig these files use: groups of
`name = value;` settings with integer, floating-point and string values, and
comments.
"""
import re
from dataclasses import dataclass, field, fields

from .network import MaturationParams, NoiseParams, SynapticParams
from .topology import TopologyParams


@dataclass
class ConnectionParams:
    """ConnectionParameters (c++/include/ConfigurationNetworkGrowth.h:97-105)."""
    Gei_max: float = 0.25
    Gie_max: float = 0.0
    delay_constant: float = 0.0


@dataclass
class GrowthConfig:
    synaptic: SynapticParams = field(default_factory=SynapticParams)
    maturation: MaturationParams = field(default_factory=MaturationParams)
    noise: NoiseParams = field(default_factory=NoiseParams)
    connection: ConnectionParams = field(default_factory=ConnectionParams)


_TOKEN = re.compile(r'\s*(?:(?P<name>[A-Za-z_][A-Za-z0-9_\-]*)|(?P<str>"(?:[^"\\]|\\.)*")|'
                    r'(?P<num>[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?L?)|(?P<punct>[={};:,]))')


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return "\n".join(re.split(r"//|#", line)[0] for line in text.splitlines())


def parse(text):
    """Groups of settings -> {group: {name: value}}; top-level scalars kept too."""
    toks, pos, text = [], 0, _strip_comments(text)
    while pos < len(text):
        if text[pos:].strip() == "":
            break
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError("cannot parse configuration near %r" % text[pos:pos + 30])
        toks.append((m.lastgroup, m.group(m.lastgroup)))
        pos = m.end()
    out, stack, i = {}, [], 0
    cur = out
    while i < len(toks):
        kind, tok = toks[i]
        if kind == "punct" and tok == "}":
            cur = stack.pop()
            i += 1
            continue
        if kind == "punct" and tok in ";,":
            i += 1
            continue
        if kind != "name" or i + 1 >= len(toks) or toks[i + 1][1] not in "=:":
            raise ValueError("expected a setting name at %r" % tok)
        name = tok
        kind, val = toks[i + 2]
        if kind == "punct" and val == "{":
            group = {}
            cur[name] = group
            stack.append(cur)
            cur = group
            i += 3
            continue
        if kind == "str":
            cur[name] = bytes(val[1:-1], "utf-8").decode("unicode_escape")
        elif kind == "num":
            v = val.rstrip("L")
            cur[name] = int(v) if re.fullmatch(r"[-+]?\d+", v) else float(v)
        elif kind == "name" and val in ("true", "false"):
            cur[name] = val == "true"
        else:
            raise ValueError("unsupported value %r for %s" % (val, name))
        i += 3
    return out


def _fill(obj, group):
    for f in fields(obj):
        if f.name in group:
            v = group[f.name]
            setattr(obj, f.name, int(v) if f.type in (int, "int") else (float(v) if not isinstance(v, str) else v))
    return obj


def read_growth_configuration(path):
    with open(path) as f:
        g = parse(f.read())
    c = GrowthConfig()
    _fill(c.synaptic, g.get("synapticParameters", {}))
    _fill(c.maturation, g.get("maturationParameters", {}))
    _fill(c.noise, g.get("noiseParameters", {}))
    _fill(c.connection, g.get("connectionParameters", {}))
    return c


def _f(x):
    return "%.6f" % x


def growth_configuration_text(c: GrowthConfig):
    s, m, n, k = c.synaptic, c.maturation, c.noise, c.connection
    L = ["synapticParameters = ", "{", "\tNss = %d;" % s.Nss, ""]
    L += ["\t%s = %s;" % (name, _f(getattr(s, name))) for name in ("A_P", "A_D", "T_0", "T_P", "T_D", "TAU_P")]
    L += ["\tTAU_D = %s;" % _f(s.TAU_D), ""]
    L += ["\tBETA = %s;" % _f(s.BETA), "\tBETA_SUPERSYNAPSE = %s;" % _f(s.BETA_SUPERSYNAPSE), ""]
    L += ["\tACTIVATION_THRESHOLD = %s;" % _f(s.ACTIVATION_THRESHOLD),
          "\tSUPERSYNAPSE_THRESHOLD = %s;" % _f(s.SUPERSYNAPSE_THRESHOLD), "\tWEIGHT_MAX = %s;" % _f(s.WEIGHT_MAX),
          "", "};", ""]
    L += ["maturationParameters = ", "{"]
    for a in ("E_GABA", "E_REST", "AD", "GK", "GNA", "GCA", "GCAK", "GSL", "GDL", "RC"):
        L += ["\t%s_MATURE = %s;" % (a, _f(getattr(m, a + "_MATURE"))),
              "\t%s_IMMATURE = %s;" % (a, _f(getattr(m, a + "_IMMATURE"))), ""]
    L += ["\tDEATH_RATE_THRESHOLD = %s;" % _f(m.DEATH_RATE_THRESHOLD), "",
          "\tRATE_WINDOW_LONG = %d;" % m.RATE_WINDOW_LONG, "", "};", ""]
    L += ["noiseParameters = ", "{"]
    L += ["\t%s = %s;" % (name, _f(getattr(n, name))) for name in
          ("white_noise_mean_soma", "white_noise_std_soma", "white_noise_mean_dend", "white_noise_std_dend")]
    L += ["};", ""]
    L += ["connectionParameters = ", "{", "\tGei_max = %s;" % _f(k.Gei_max), "\tGie_max = %s;" % _f(k.Gie_max),
          "\tdelay_constant = %s;" % _f(k.delay_constant), "", "};", ""]
    return "\n".join(L) + "\n"


def write_growth_configuration(path, c: GrowthConfig):
    with open(path, "w") as f:
        f.write(growth_configuration_text(c))


def topology_configuration_text(p: TopologyParams):
    return ("spatialParameters = \n{\n\tarrangement = \"%s\";\n\tA_RA2I = %s;\n\tSIGMA_RA2I = %s;\n"
            "\tB_I2RA = %s;\n\tSIGMA_I2RA = %s;\n\n};\n\n") % (p.arrangement, _f(p.A_RA2I), _f(p.SIGMA_RA2I),
                                                             _f(p.B_I2RA), _f(p.SIGMA_I2RA))


def write_topology_configuration(path, p: TopologyParams):
    with open(path, "w") as f:
        f.write(topology_configuration_text(p))


def read_topology_configuration(path, connection: ConnectionParams = None):
    """spatialParameters; the connection parameters (weights, delays) live in the
    growth configuration and are merged in when given."""
    with open(path) as f:
        g = parse(f.read()).get("spatialParameters", {})
    p = TopologyParams()
    for name in ("A_RA2I", "SIGMA_RA2I", "B_I2RA", "SIGMA_I2RA"):
        if name in g:
            setattr(p, name, float(g[name]))
    if "arrangement" in g:
        p.arrangement = g["arrangement"]
    if connection is not None:
        p.Gei_max, p.Gie_max, p.delay_constant = connection.Gei_max, connection.Gie_max, connection.delay_constant
    return p

"""Algorithmic work per neuron-step, counted by evaluating the engine's stepper
expression trees (csrc/hvc_neuron.hip.h) on an operation-counting float.

The counts are the floating-point operations one neuron-step performs:
additions, subtractions, multiplications and divisions of the stage and RHS
evaluations, and calls of exp / log.  The conductance decay factors shared by
all neurons come from a per-step table and are not counted per neuron.
``FLOPS_PER_EXP`` is the operation count of one exp in each precision form:
FP64 (precision 0) 20, the engine's own exp (hvc_math.hip.h: 1 mul for the
reduction argument, rint, 2 fma reduction, 1 mul r^2, 5 fma polynomial, 1 fma
+ 1 add reconstruction, 1 mul scale; fma = 2 flops); FP32 (precision 1) 1,
the hardware exp2 (v_exp_f32) -- the argument's scaling by log2(e) / B is the
expression's division by B, counted with the divisions.

``flops_per_step(kind, model, precision)`` is what bench.py uses for the roofline.
"""
import math

FLOPS_PER_EXP = {0: 20, 1: 1}


class Counter:
    def __init__(self):
        self.n = {"add": 0, "mul": 0, "div": 0, "exp": 0}

    def flops(self, precision=0):
        return self.n["add"] + self.n["mul"] + self.n["div"] + FLOPS_PER_EXP[precision] * self.n["exp"]


class CF:
    """float that counts its arithmetic in a shared Counter."""
    __slots__ = ("v", "c")

    def __init__(self, v, c):
        self.v = float(v)
        self.c = c

    def _w(self, x):
        return x.v if isinstance(x, CF) else float(x)

    def __add__(self, o):
        self.c.n["add"] += 1
        return CF(self.v + self._w(o), self.c)

    __radd__ = __add__

    def __sub__(self, o):
        self.c.n["add"] += 1
        return CF(self.v - self._w(o), self.c)

    def __rsub__(self, o):
        self.c.n["add"] += 1
        return CF(self._w(o) - self.v, self.c)

    def __mul__(self, o):
        self.c.n["mul"] += 1
        return CF(self.v * self._w(o), self.c)

    __rmul__ = __mul__

    def __truediv__(self, o):
        self.c.n["div"] += 1
        return CF(self.v / self._w(o), self.c)

    def __rtruediv__(self, o):
        self.c.n["div"] += 1
        return CF(self._w(o) / self.v, self.c)

    def __neg__(self):
        return CF(-self.v, self.c)


def _exp(x):
    x.c.n["exp"] += 1
    return CF(math.exp(x.v), x.c)


# ---------------- HVC-RA (ra_rhs / ra_debraband / ra_runge4) ----------------
def _ra_rhs(model, s, EL, GCa, mus, mud, vs, vd, n, h, l, r, c, ca, E):
    m = 1 / (1 + _exp(-(vs + 30) / 9.5))
    m3 = m * m * m
    n4 = n * n * n * n
    acc = -0.1 * (vs - EL) - 60.0 * m3 * h * (vs - 55.0) - 8.0 * n4 * (vs + 90.0)
    if model == 0:
        acc = acc - 3.5 * l * (vs + 90.0)
    kVs = acc - (s["Gi_s"] * E) * (vs + 80.0) - (s["Ge_s"] * E) * vs + mus + 100000 * (vd - vs) / 275000.0
    r2 = r * r
    kVd = (-0.1 * (vd - EL) - GCa * r2 * (vd - 120.0) - 150.0 * (vd + 90.0) * c / (1 + 6 / ca)
           - (s["Ge_d"] * E) * vd - (s["Gi_d"] * E) * (vd + 80.0) + mud + 100000 * (vs - vd) / 550000.0)
    kn = (1 / (1 + _exp(-(vs + 35) / 10)) - n) / (0.1 + 0.5 / (1 + _exp((vs + 27) / 15)))
    kl = (1 / (1 + _exp(-(vs + 40) / 5)) - l) / 10.0 if model == 0 else 0.0
    kh = (1 / (1 + _exp((vs + 45) / 7)) - h) / (0.1 + 0.75 / (1 + _exp((vs + 40.5) / 6)))
    kr = (1 / (1 + _exp(-(vd + 5) / 10)) - r)
    kc = (1 / (1 + _exp(-(vd - 10) / 7)) - c) / (15.0 if model == 0 else 10.0)
    kCa = (-0.1 * GCa * r * r * (vd - 120.0) - 0.02 * ca)
    return dict(Vs=kVs, Vd=kVd, n=kn, l=kl, h=kh, r=kr, c=kc, Ca=kCa)


def _ra_state(cnt):
    vals = dict(Vs=-79.97619025, Vd=-79.97268759, n=0.01101284, l=0.0, h=0.9932845, r=0.00055429,
                c=0.00000261762353, Ca=0.01689572, Ge_s=0.0, Gi_s=0.0, Ge_d=0.5, Gi_d=0.2)
    return {k: CF(v, cnt) for k, v in vals.items()}


def count_ra_step(model):
    cnt = Counter()
    s = _ra_state(cnt)
    dt = 0.02
    EL, GCa, mus, mud = -80.0, 55.0, 0.0, 0.0
    E1, Eh, E13, E23 = 0.996, 0.998, 0.9987, 0.9973
    keys = ("Vs", "Vd", "n", "h", "l", "r", "c", "Ca")

    def rhs(x, E):
        return _ra_rhs(model, s, EL, GCa, mus, mud, x["Vs"], x["Vd"], x["n"], x["h"], x["l"], x["r"], x["c"],
                       x["Ca"], E)

    if model == 0:
        sq_s, sq_d = 0.0141, 0.0283
        J1s, J2s, J1d, J2d = 1.0, -1.0, 2.449489742783178, 0.0
        x1 = dict(s)
        x1["Vs"] = s["Vs"] + sq_s * (-0.0184 * J1s - 0.1866 * J2s) / 5000.0
        x1["Vd"] = s["Vd"] + sq_d * (-0.0184 * J1d - 0.1866 * J2d) / 10000.0
        a1 = rhs(x1, 1.0)
        x2 = {k: s[k] + dt * a1[k] for k in keys}
        x2["Vs"] = x2["Vs"] + sq_s * (0.80 * J1s - 0.86 * J2s) / 5000.0
        x2["Vd"] = x2["Vd"] + sq_d * (0.80 * J1d - 0.86 * J2d) / 10000.0
        a2 = rhs(x2, E1)
        x3 = {k: s[k] + dt * (3 * a1[k] + a2[k]) / 8 for k in keys}
        x3["Vs"] = x3["Vs"] + sq_s * (0.51 * J1s - 0.47 * J2s) / 5000.0
        x3["Vd"] = x3["Vd"] + sq_d * (0.51 * J1d - 0.47 * J2d) / 10000.0
        a3 = rhs(x3, Eh)
        x4 = {k: s[k] + dt * (-0.45 * a1[k] - 0.48 * a2[k] + 1.94 * a3[k]) for k in keys}
        x4["Vs"] = x4["Vs"] + sq_s * (0.98 * J1s + 0.31 * J2s) / 5000.0
        x4["Vd"] = x4["Vd"] + sq_d * (0.98 * J1d + 0.31 * J2d) / 10000.0
        a4 = rhs(x4, E1)
        for k in keys:
            s[k] = s[k] + dt * (a1[k] / 6 - 0.0054 * a2[k] + 2 * a3[k] / 3 + 0.172 * a4[k])
        s["Vs"] = s["Vs"] + sq_s * J1s * 100000 / 5000.0
        s["Vd"] = s["Vd"] + sq_d * J1d * 100000 / 10000.0
        for g in ("Ge_d", "Ge_s", "Gi_d", "Gi_s"):
            s[g] = s[g] * E1
    else:
        keys = ("Vs", "Vd", "n", "h", "r", "c", "Ca")
        k1 = rhs(s, 1.0)
        k2 = rhs({**s, **{k: s[k] + dt * k1[k] / 3 for k in keys}}, E13)
        k3 = rhs({**s, **{k: s[k] + dt * (-k1[k] / 3 + k2[k]) for k in keys}}, E23)
        k4 = rhs({**s, **{k: s[k] + dt * (k1[k] - k2[k] + k3[k]) for k in keys}}, E1)
        for g in ("Ge_d", "Ge_s", "Gi_d", "Gi_s"):
            s[g] = s[g] * E1
        for k in keys:
            s[k] = s[k] + dt * (k1[k] + 3 * k2[k] + 3 * k3[k] + k4[k]) / 8
    return cnt


# ---------------- HVC-I (i_rhs / i_dp8 / i_runge4) ----------------
def _i_rhs(s, v, n1, m1, h1, w1, e2, e5):
    n4 = n1 * n1 * n1 * n1
    m3 = m1 * m1 * m1
    kV = (-0.1 * (v + 65.0) - 100.0 * h1 * m3 * (v - 55.0) - 20.0 * n4 * (v + 80.0) - 500.0 * w1 * (v + 80.0)
          - (s["Ge"] * e2) * v - (s["Gi"] * e5) * (v + 75.0) + 100000 * 0.0 / 6000.0)
    an = 0.15 * (v + 15) / (1 - _exp(-(v + 15) / 10))
    bn = 0.2 * _exp(-(v + 25) / 80)
    am = (v + 22) / (1 - _exp(-(v + 22) / 10))
    bm = 40 * _exp(-(v + 47) / 18)
    ah = 0.7 * _exp(-(v + 34) / 20)
    bh = 10 / (1 + _exp(-(v + 4) / 10))
    wI = 1 / (1 + _exp(-v / 5))
    return dict(V=kV, n=an * (1 - n1) - bn * n1, m=am * (1 - m1) - bm * m1, h=ah * (1 - h1) - bh * h1,
                w=(wI - w1))


# DP8 stage coefficients as (numerator, denominator) pairs, HHI_buffer.cpp:452-669
_DP8 = [
    [(1, 18)],
    [(1, 48), (3, 48)],
    [(1, 32), (0, 1), (3, 32)],
    [(20, 64), (0, 1), (-75, 64), (75, 64)],
    [(3, 80), (0, 1), (0, 1), (15, 80), (12, 80)],
    [(29443841, 614563906), None, None, (77736538, 692538347), (-28693883, 1125000000), (23124283, 1800000000)],
    [(16016141, 946692911), None, None, (61564180, 158732637), (22789713, 633445777), (545815736, 2771057229),
     (-180193667, 1043307555)],
    [(39632708, 573591083), None, None, (-433636366, 683701615), (-421739975, 2616292301),
     (100302831, 723423059), (790204164, 839813087), (800635310, 3783071287)],
    [(246121993, 1340847787), None, None, (-37695042795, 15268766246), (-309121744, 1061227803),
     (-12992083, 490766935), (6005943493, 2108947869), (393006217, 1396673457), (123872331, 1001029789)],
    [(-1028468189, 846180014), None, None, (8478235783, 508512852), (1311729495, 1432422823),
     (-10304129995, 1701304382), (-48777925059, 3047939560), (15336726248, 1032824649),
     (-45442868181, 3398467696), (3065993473, 597172653)],
    [(185892177, 718116043), None, None, (-3185094517, 667107341), (-477755414, 1098053517),
     (-703635378, 230739211), (5731566787, 1027545527), (5232866602, 850066563), (-4093664535, 808688257),
     (3962137247, 1805957418), (65686358, 487910083)],
]
_DP8_FINAL = [(13451932, 455176623), None, None, None, None, (-808719846, 976000145),
              (1757004468, 5645159321), (656045339, 265891186), (-3867574721, 1518517206),
              (465885868, 322736535), (53011238, 667516719), (2, 45)]


def count_i_step(model):
    cnt = Counter()
    s = {k: CF(v, cnt) for k, v in dict(V=-66.0, n=0.125, m=0.0, h=0.99, w=0.0, Ge=0.5, Gi=0.2).items()}
    keys = ("V", "n", "m", "h", "w")
    dt = 0.02
    ks = [_i_rhs(s, s["V"], s["n"], s["m"], s["h"], s["w"], 1.0, 1.0)]
    if model == 0:
        for st, row in enumerate(_DP8):
            x = {}
            for k in keys:
                if st < 5:
                    # stages 2-6 sum small integer multiples, then divide once (HHI_buffer.cpp:452-516)
                    terms = [(a, ks[j][k]) for j, cd in enumerate(row) if cd and cd[0] != 0 for a in [cd[0]]]
                    acc = None
                    for a, kk in terms:
                        t = kk if abs(a) == 1 else abs(a) * kk
                        acc = t if acc is None else (acc - t if a < 0 else acc + t)
                    x[k] = s[k] + dt * acc / row[0][1]
                else:
                    acc = None
                    for j, cd in enumerate(row):
                        if cd is None:
                            continue
                        t = abs(cd[0]) * ks[j][k] / cd[1]
                        acc = t if acc is None else (acc - t if cd[0] < 0 else acc + t)
                    x[k] = s[k] + dt * acc
            ks.append(_i_rhs(s, x["V"], x["n"], x["m"], x["h"], x["w"], 0.99, 0.995))
        for k in keys:
            acc = None
            for j, cd in enumerate(_DP8_FINAL):
                if cd is None:
                    continue
                t = abs(cd[0]) * ks[j][k] / cd[1]
                acc = t if acc is None else (acc - t if cd[0] < 0 else acc + t)
            s[k] = s[k] + dt * acc
    else:
        k1 = ks[0]
        k2 = _i_rhs(s, *[s[k] + dt * k1[k] / 3 for k in keys], 0.99, 0.995)
        k3 = _i_rhs(s, *[s[k] + dt * (-k1[k] / 3 + k2[k]) for k in keys], 0.99, 0.995)
        k4 = _i_rhs(s, *[s[k] + dt * (k1[k] - k2[k] + k3[k]) for k in keys], 0.99, 0.995)
        for k in keys:
            s[k] = s[k] + dt * (k1[k] + 3 * k2[k] + 3 * k3[k] + k4[k]) / 8
    s["Ge"] = s["Ge"] * 0.99
    s["Gi"] = s["Gi"] * 0.995
    return cnt


def flops_per_step(kind, model, precision=0):
    """Operations of one neuron-step in a precision form (0 FP64, 1 FP32); kind 'ra' or 'i'."""
    cnt = count_ra_step(model) if kind == "ra" else count_i_step(model)
    return cnt.flops(precision)


if __name__ == "__main__":
    for kind, model, name in (("ra", 0, "RA Debraband"), ("ra", 1, "RA RK4"), ("i", 0, "I DP8"), ("i", 1, "I RK4")):
        c = count_ra_step(model) if kind == "ra" else count_i_step(model)
        print("%-13s %s flops FP64=%d FP32=%d" % (name, c.n, c.flops(0), c.flops(1)))

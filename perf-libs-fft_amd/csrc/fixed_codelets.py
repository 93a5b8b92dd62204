"""Op lists of the fixed-point codelets, and the generator of
fixed_codelets.hpp.

The reference builds its fixed-point radix-2/4/8/16 codelets from symbolic
expression lists: Winograd "separated" forms for 2, 4 and 8
(src/providers/jit/winograd_z{2,4,8}_separated.cpp, driven by
src/providers/jit/wfta_nfact.cpp:22-171) and, for fixed point, the split-radix
16 (use_alternative, wfta_nfact.cpp:89-93; split_radix_z16.cpp:13-195). The
lists below restate those expressions in this module's own notation:

    (dst, "+", a, b)        saturating add        (emit.hpp: SQADD)
    (dst, "-", a, b)        saturating subtract   (SQSUB)
    (dst, "=", a)           copy (no arithmetic)
    (dst, "R", a, c)        a times the real constant c
    (dst, "I", a, s)        a times the imaginary constant i*s

and lower() turns them into the fixed-point operations the reference emits
(print_algo.cpp:2186-2290 lowering, do_opt rewrites, emit.hpp:150-430):

    "R" with c == 1.0       folded away (do_opt: multiply by uniform 1)
    "R"                     MUL(a, (Q(c), Q(c)))   per-lane rounding multiply
    "I"                     MUL(rev(a), (Q(-s), Q(s)))   rev = swap lanes
    "+"/"-" of two MULs     MAC: one widening accumulate, one rounding narrow
    otherwise               saturating add / subtract

Q(c) = sat(trunc(float(c) * 2^B)) (src/plfft_convert.hpp:40-62).
Run as a script to regenerate fixed_codelets.hpp.
"""
import math
import os
import sys

# ---- radix 2 (winograd_z2_separated.cpp): in, multiply by 1 (folded), out
Z2 = [("y0", "+", "x0", "x1"), ("y1", "-", "x0", "x1")]

# ---- radix 4 (winograd_z4_separated.cpp)
Z4 = [
    ("l0", "+", "x0", "x2"), ("l1", "-", "x0", "x2"),
    ("l2", "+", "x1", "x3"), ("l3", "-", "x1", "x3"),
    ("l4", "+", "l0", "l2"), ("l5", "-", "l0", "l2"),
    # S = [l4, l5, l1, l3]; multipliers [1, 1, 1, i]
    ("m3", "I", "l3", 1.0),
    # out: S[0] = X0, S[1] = X2 - X3, S[2] = X1, S[3] = X2 + X3
    ("o0", "+", "l1", "m3"), ("o1", "-", "l1", "m3"),
    ("y0", "=", "l4"), ("y1", "=", "o1"), ("y2", "=", "l5"), ("y3", "=", "o0"),
]

# ---- radix 8 (winograd_z8_separated.cpp); u = 2 pi / 8, constants as the
# generator computes them (std::sin / std::cos in double)
_U = 2.0 * math.pi / 8.0
Z8 = [
    ("l0", "+", "x0", "x4"), ("l1", "-", "x0", "x4"),
    ("l2", "+", "x2", "x6"), ("l3", "-", "x2", "x6"),
    ("l4", "+", "x1", "x5"), ("l5", "-", "x1", "x5"),
    ("l6", "+", "x3", "x7"), ("l7", "-", "x3", "x7"),
    ("l8", "+", "l0", "l2"), ("l9", "-", "l0", "l2"),
    ("l10", "+", "l4", "l6"), ("l11", "-", "l4", "l6"),
    ("l12", "+", "l8", "l10"), ("l13", "-", "l8", "l10"),
    ("l14", "+", "l5", "l7"), ("l15", "-", "l5", "l7"),
    # S = [l12, l13, l9, l11, l1, l3, l14, l15]
    # multipliers [1, 1, 1, i sin 2u, 1, i sin 2u, i sin u, cos u]
    ("m3", "I", "l11", math.sin(2.0 * _U)),
    ("m5", "I", "l3", math.sin(2.0 * _U)),
    ("m6", "I", "l14", math.sin(_U)),
    ("m7", "R", "l15", math.cos(_U)),
    # out (X = [S0, S1, S2, m3, S4, m5, m6, m7])
    ("o0", "+", "l9", "m3"), ("o1", "-", "l9", "m3"),
    ("o2", "+", "l1", "m7"), ("o3", "-", "l1", "m7"),
    ("o4", "+", "m5", "m6"), ("o5", "-", "m5", "m6"),
    ("o6", "+", "o2", "o4"), ("o7", "-", "o2", "o4"),
    ("o8", "+", "o3", "o5"), ("o9", "-", "o3", "o5"),
    ("y0", "=", "l12"), ("y1", "=", "o7"), ("y2", "=", "o1"), ("y3", "=", "o8"),
    ("y4", "=", "l13"), ("y5", "=", "o9"), ("y6", "=", "o0"), ("y7", "=", "o6"),
]

# ---- radix 16, split radix (split_radix_z16.cpp); Y are temporaries and,
# at the end, the outputs
_C1 = 0.707106781186547572737310929369
_C1N = -0.707106781186547461715008466854
Z16 = [
    ("Y0", "+", "x0", "x8"), ("Y1", "-", "x0", "x8"), ("Y2", "=", "x4"), ("Y3", "=", "x12"),
    ("l0", "+", "Y2", "Y3"), ("l1", "-", "Y2", "Y3"), ("l2", "I", "l1", 1.0),
    ("l3", "+", "Y0", "l0"), ("l4", "-", "Y0", "l0"), ("l5", "-", "Y1", "l2"),
    ("l6", "+", "Y1", "l2"),
    ("Y0", "=", "l3"), ("Y1", "=", "l5"), ("Y2", "=", "l4"), ("Y3", "=", "l6"),
    ("Y4", "+", "x2", "x10"), ("Y5", "-", "x2", "x10"), ("Y6", "+", "x6", "x14"),
    ("Y7", "-", "x6", "x14"),
    ("l7", "+", "Y4", "Y6"), ("l8", "-", "Y4", "Y6"), ("l9", "I", "l8", 1.0),
    ("l10", "+", "Y0", "l7"), ("l11", "-", "Y0", "l7"), ("l12", "-", "Y2", "l9"),
    ("l13", "+", "Y2", "l9"),
    ("Y0", "=", "l10"), ("Y2", "=", "l12"), ("Y4", "=", "l11"), ("Y6", "=", "l13"),
    ("l14", "R", "Y5", _C1), ("l15", "I", "l14", -1.0),
    ("l16", "R", "Y7", _C1N), ("l17", "I", "l16", 1.0),
    ("l18", "+", "l14", "l15"), ("l19", "+", "l16", "l17"),
    ("l20", "+", "l18", "l19"), ("l21", "-", "l19", "l18"), ("l22", "I", "l21", -1.0),
    ("l23", "+", "Y1", "l20"), ("l24", "-", "Y1", "l20"), ("l25", "-", "Y3", "l22"),
    ("l26", "+", "Y3", "l22"),
    ("Y1", "=", "l23"), ("Y3", "=", "l25"), ("Y5", "=", "l24"), ("Y7", "=", "l26"),
    ("Y8", "+", "x1", "x9"), ("Y9", "-", "x1", "x9"), ("Y10", "=", "x5"), ("Y11", "=", "x13"),
    ("l27", "+", "Y10", "Y11"), ("l28", "-", "Y10", "Y11"), ("l29", "I", "l28", 1.0),
    ("l30", "+", "Y8", "l27"), ("l31", "-", "Y8", "l27"), ("l32", "-", "Y9", "l29"),
    ("l33", "+", "Y9", "l29"),
    ("Y8", "=", "l30"), ("Y9", "=", "l32"), ("Y10", "=", "l31"), ("Y11", "=", "l33"),
    ("Y12", "+", "x3", "x11"), ("Y13", "-", "x3", "x11"), ("Y14", "=", "x7"),
    ("Y15", "=", "x15"),
    ("l34", "+", "Y14", "Y15"), ("l35", "-", "Y14", "Y15"), ("l36", "I", "l35", 1.0),
    ("l37", "+", "Y12", "l34"), ("l38", "-", "Y12", "l34"), ("l39", "-", "Y13", "l36"),
    ("l40", "+", "Y13", "l36"),
    ("Y12", "=", "l37"), ("Y13", "=", "l39"), ("Y14", "=", "l38"), ("Y15", "=", "l40"),
    ("l41", "+", "Y8", "Y12"), ("l42", "-", "Y8", "Y12"), ("l43", "I", "l42", 1.0),
    ("l44", "+", "Y0", "l41"), ("l45", "-", "Y0", "l41"), ("l46", "-", "Y4", "l43"),
    ("l47", "+", "Y4", "l43"),
    ("Y0", "=", "l44"), ("Y4", "=", "l46"), ("Y8", "=", "l45"), ("Y12", "=", "l47"),
    ("l48", "R", "Y9", 0.923879532511286738483136105060),
    ("l49", "I", "Y9", -0.382683432365089781779232680492),
    ("l50", "R", "Y13", 0.382683432365089837290383911750),
    ("l51", "I", "Y13", -0.923879532511286738483136105060),
    ("l52", "+", "l48", "l49"), ("l53", "+", "l50", "l51"), ("l54", "+", "l52", "l53"),
    ("l55", "-", "l53", "l52"), ("l56", "I", "l55", -1.0),
    ("l57", "+", "Y1", "l54"), ("l58", "-", "Y1", "l54"), ("l59", "-", "Y5", "l56"),
    ("l60", "+", "Y5", "l56"),
    ("Y1", "=", "l57"), ("Y5", "=", "l59"), ("Y9", "=", "l58"), ("Y13", "=", "l60"),
    ("l61", "R", "Y10", _C1), ("l62", "I", "l61", -1.0),
    ("l63", "R", "Y14", _C1N), ("l64", "I", "l63", 1.0),
    ("l65", "+", "l61", "l62"), ("l66", "+", "l63", "l64"), ("l67", "+", "l65", "l66"),
    ("l68", "-", "l66", "l65"), ("l69", "I", "l68", -1.0),
    ("l70", "+", "Y2", "l67"), ("l71", "-", "Y2", "l67"), ("l72", "-", "Y6", "l69"),
    ("l73", "+", "Y6", "l69"),
    ("Y2", "=", "l70"), ("Y6", "=", "l72"), ("Y10", "=", "l71"), ("Y14", "=", "l73"),
    ("l74", "R", "Y11", 0.382683432365089837290383911750),
    ("l75", "I", "Y11", -0.923879532511286738483136105060),
    ("l76", "R", "Y15", -0.923879532511286849505438567576),
    ("l77", "I", "Y15", 0.382683432365089670756930217976),
    ("l78", "+", "l74", "l75"), ("l79", "+", "l76", "l77"), ("l80", "+", "l78", "l79"),
    ("l81", "-", "l79", "l78"), ("l82", "I", "l81", -1.0),
    ("l83", "+", "Y3", "l80"), ("l84", "-", "Y3", "l80"), ("l85", "-", "Y7", "l82"),
    ("l86", "+", "Y7", "l82"),
    ("Y3", "=", "l83"), ("Y7", "=", "l85"), ("Y11", "=", "l84"), ("Y15", "=", "l86"),
] + [(f"y{k}", "=", f"Y{k}") for k in range(16)]

CODELETS = {2: Z2, 4: Z4, 8: Z8, 16: Z16}


def eval_float(ops, x):
    """Evaluate an op list in complex floating point (for checking)."""
    env = {f"x{i}": complex(v) for i, v in enumerate(x)}
    for op in ops:
        d, kind = op[0], op[1]
        if kind == "+":
            env[d] = env[op[2]] + env[op[3]]
        elif kind == "-":
            env[d] = env[op[2]] - env[op[3]]
        elif kind == "=":
            env[d] = env[op[2]]
        elif kind == "R":
            env[d] = env[op[2]] * op[3]
        else:
            env[d] = env[op[2]] * 1j * op[3]
    return [env[f"y{k}"] for k in range(len(x))]


def q_const(c, bits):
    """sat(trunc(float(c) * 2^bits)) with float = IEEE binary32."""
    import struct
    f = struct.unpack("f", struct.pack("f", c))[0]
    v = int(f * (1 << bits))  # exact scaling, truncation toward zero
    return max(-(1 << bits), min((1 << bits) - 1, v))


def lower(ops, r=None):
    """Fixed-point lowering. Returns a list of
      ("add"|"sub", d, a, b)
      ("mul", d, a, rev, c0, c1)        d = a' * (c0, c1), a' = rev(a) if rev
      ("mac", d, sign, (a, rev, c0, c1), (b, rev, c0, c1))
      ("copy", d, a)
    with constants as doubles (converted per bit width by the emitters)."""
    out = []
    alias = {}
    muls = {}  # value -> (operand, rev, c0, c1) of the multiply that made it

    def res(v):
        return alias.get(v, v)

    for op in ops:
        d, kind = op[0], op[1]
        if kind == "=":
            alias[d] = res(op[2])
            continue
        alias.pop(d, None)
        if kind == "R":
            a = res(op[2])
            if op[3] == 1.0:  # do_opt: multiply by a uniform 1 is folded
                alias[d] = a
                continue
            out.append(("mul", d, a, False, op[3], op[3]))
            muls[d] = (a, False, op[3], op[3])
        elif kind == "I":
            a = res(op[2])
            out.append(("mul", d, a, True, -op[3], op[3]))
            muls[d] = (a, True, -op[3], op[3])
        else:
            a, b = res(op[2]), res(op[3])
            if a in muls and b in muls:
                out.append(("mac", d, 1 if kind == "+" else -1, muls[a], muls[b]))
            else:
                out.append(("add" if kind == "+" else "sub", d, a, b))
        if kind in "+-":
            muls.pop(d, None)
    if r is None:
        r = 1 + max(int(o[0][1:]) for o in ops if o[0].startswith("y"))
    outs = [res(f"y{k}") for k in range(r)]
    return out, outs


def check():
    import cmath
    for r, ops in CODELETS.items():
        x = [complex(math.cos(0.3 * i + 1.1) * (i + 1), math.sin(0.7 * i) - 0.2 * i)
             for i in range(r)]
        y = eval_float(ops, x)
        ref = [sum(x[j] * cmath.exp(-2j * math.pi * j * k / r) for j in range(r))
               for k in range(r)]
        err = max(abs(a - b) for a, b in zip(y, ref))
        assert err < 1e-9, (r, err)
    return True


def emit():
    """fixed_codelets.hpp: one device function per radix over an Ops policy
    (add, sub, mul<rev>(a, q0, q1), mac<sign, reva, revb>(a, qa0, qa1, b,
    qb0, qb1)) with the constants converted for B bits at compile time."""
    lines = [
        "// fixed_codelets.hpp -- generated by fixed_codelets.py (do not edit):",
        "// the reference's fixed-point radix-2/4/8/16 codelet op lists, lowered",
        "// to saturating add/sub, per-lane rounding multiplies and one-rounding",
        "// multiply-accumulates (see fixed_codelets.py for the sources).",
        "#pragma once",
        "",
        "namespace plf {",
        "",
        "// Q(c) = sat(trunc(float(c) * 2^B))",
        "template <int B>",
        "constexpr int fxq(double c) {",
        "  const double v = (double)(float)c * (double)(1 << B);",
        "  const long t = (long)v;  // truncation toward zero",
        "  return t > (1 << B) - 1 ? (1 << B) - 1 : (t < -(1 << B) ? -(1 << B) : (int)t);",
        "}",
        "",
    ]
    for r, ops in CODELETS.items():
        low, outs = lower(ops)
        # drop values used by nothing (products consumed only through a
        # fused multiply-accumulate)
        used = set(outs)
        for op in reversed(low):
            if op[1] not in used:
                continue
            if op[0] in ("add", "sub"):
                used.update(op[2:4])
            elif op[0] == "mul":
                used.add(op[2])
            else:
                used.update((op[3][0], op[4][0]))
        low = [op for op in low if op[1] in used]
        lines.append(f"// forward DFT_{r} on a[0..{r}) in place, natural order")
        lines.append("template <class O, int B>")
        lines.append(f"__device__ __forceinline__ void fx_dft{r}(typename O::V *a) {{")
        lines.append("  using V = typename O::V;")
        for i in range(r):
            lines.append(f"  const V x{i} = a[{i}];")
        for op in low:
            if op[0] in ("add", "sub"):
                lines.append(f"  const V {op[1]} = O::{op[0]}({op[2]}, {op[3]});")
            elif op[0] == "mul":
                _, d, a, rv, c0, c1 = op
                lines.append(f"  const V {d} = O::template mul<{str(rv).lower()}>({a}, "
                             f"fxq<B>({c0!r}), fxq<B>({c1!r}));")
            else:
                _, d, sg, (a, ra, a0, a1), (b, rb, b0, b1) = op
                lines.append(f"  const V {d} = O::template mac<{sg}, {str(ra).lower()}, "
                             f"{str(rb).lower()}>({a}, fxq<B>({a0!r}), fxq<B>({a1!r}), "
                             f"{b}, fxq<B>({b0!r}), fxq<B>({b1!r}));")
        for k, o in enumerate(outs):
            lines.append(f"  a[{k}] = {o};")
        lines.append("}")
        lines.append("")
    lines.append("}  // namespace plf")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    check()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixed_codelets.hpp")
    with open(out, "w") as f:
        f.write(emit())
    print("wrote", out, file=sys.stderr)

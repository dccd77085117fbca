"""Explicit Runge-Kutta tables of ReactorArkode (ode.rk_method,
Source/Reactions/ReactorArkode.cpp:26-83), as exact rationals.

Each entry: (ARKODE name, c, A (strictly lower rows 2..s), b, d, q, p) with
b the weights of the order-q solution and d those of the order-p embedding
that estimates its error.  The values are the published coefficients of each
method (Heun-Euler; Bogacki & Shampine 1989; Kennedy & Carpenter 2003,
ARK3(2)4L[2]SA, ARK4(3)6L[2]SA and ARK5(4)8L[2]SA explicit tables; Zonneveld
1964; Sayfy & Aburub; Cash & Karp 1990; Fehlberg 1969 4(5) and 7(8);
Dormand & Prince 1980; Verner 1978 6(5)); the Kennedy-Carpenter entries are
rational approximations that meet the order conditions to ~1e-26.
``tests/test_erk_tables.py`` checks every table against all rooted-tree order
conditions up to its order (and that it fails the next one).

Codes 43 (ARK437L2SA_ERK_7_3_4) and 54 (ARK548L2SAb_ERK_8_4_5) are not
provided.

``python -m pelephysics_amd.erk_tables`` writes ``csrc/erk_tables.h``, the
same tables as correctly rounded doubles for the library.
"""
from __future__ import annotations

import os
from fractions import Fraction as F

DEFAULT_METHOD = 40  # ARKODE_ZONNEVELD_5_3_4 (ReactorArkode.H:135)
MAX_STAGES = 13


def _ark324():
    c = [0, F(1767732205903, 2027836641118), F(3, 5), 1]
    A = [[F(1767732205903, 2027836641118)],
         [F(5535828885825, 10492691773637), F(788022342437, 10882634858940)],
         [F(6485989280629, 16251701735622), F(-4246266847089, 9704473918619), F(10755448449292, 10357097424841)]]
    b = [F(1471266399579, 7840856788654), F(-4482444167858, 7529755066697), F(11266239266428, 11593286722821),
         F(1767732205903, 4055673282236)]
    d = [F(2756255671327, 12835298489170), F(-10771552573575, 22201958757719), F(9247589265047, 10645013368117),
         F(2193209047091, 5459859503100)]
    return "ARKODE_ARK324L2SA_ERK_4_2_3", c, A, b, d, 3, 2


def _ark436():
    c = [0, F(1, 2), F(83, 250), F(31, 50), F(17, 20), 1]
    A = [[F(1, 2)], [F(13861, 62500), F(6889, 62500)],
         [F(-116923316275, 2393684061468), F(-2731218467317, 15368042101831), F(9408046702089, 11113171139209)],
         [F(-451086348788, 2902428689909), F(-2682348792572, 7519795681897), F(12662868775082, 11960479115383),
          F(3355817975965, 11060851509271)],
         [F(647845179188, 3216320057751), F(73281519250, 8382639484533), F(552539513391, 3454668386233),
          F(3354512671639, 8306763924573), F(4040, 17871)]]
    b = [F(82889, 524892), 0, F(15625, 83664), F(69875, 102672), F(-2260, 8211), F(1, 4)]
    d = [F(4586570599, 29645900160), 0, F(178811875, 945068544), F(814220225, 1159782912), F(-3700637, 11593932),
         F(61727, 225920)]
    return "ARKODE_ARK436L2SA_ERK_6_3_4", c, A, b, d, 4, 3


def _ark548():
    c = [0, F(41, 100), F(2935347310677, 11292855782101), F(1426016391358, 7196633302097), F(92, 100),
         F(24, 100), F(3, 5), 1]
    A = [[F(41, 100)],
         [F(367902744464, 2072280473677), F(677623207551, 8224143866563)],
         [F(1268023523408, 10340822734521), 0, F(1029933939417, 13636558850479)],
         [F(14463281900351, 6315353703477), 0, F(66114435211212, 5879490589093), F(-54053170152839, 4284798021562)],
         [F(14090043504691, 34967701212078), 0, F(15191511035443, 11219624916014),
          F(-18461159152457, 12425892160975), F(-281667163811, 9011619295870)],
         [F(19230459214898, 13134317526959), 0, F(21275331358303, 2942455364971),
          F(-38145345988419, 4862620318723), F(-1, 8), F(-1, 8)],
         [F(-19977161125411, 11928030595625), 0, F(-40795976796054, 6384907823539),
          F(177454434618887, 12078138498510), F(782672205425, 8267701900261), F(-69563011059811, 9646580694205),
          F(7356628210526, 4942186776405)]]
    b = [F(-872700587467, 9133579230613), 0, 0, F(22348218063261, 9555858737531), F(-1143369518992, 8141816002931),
         F(-39379526789629, 19018526304540), F(32727382324388, 42900044865799), F(41, 200)]
    d = [F(-975461918565, 9796059967033), 0, 0, F(78070527104295, 32432590147079), F(-548382580838, 3424219808633),
         F(-33438840321285, 15594753105479), F(3629800801594, 4656183773603), F(4035322873751, 18575991585200)]
    return "ARKODE_ARK548L2SA_ERK_8_4_5", c, A, b, d, 5, 4


def _fehlberg78():
    c = [0, F(2, 27), F(1, 9), F(1, 6), F(5, 12), F(1, 2), F(5, 6), F(1, 6), F(2, 3), F(1, 3), 1, 0, 1]
    A = [[F(2, 27)], [F(1, 36), F(1, 12)], [F(1, 24), 0, F(1, 8)], [F(5, 12), 0, F(-25, 16), F(25, 16)],
         [F(1, 20), 0, 0, F(1, 4), F(1, 5)], [F(-25, 108), 0, 0, F(125, 108), F(-65, 27), F(125, 54)],
         [F(31, 300), 0, 0, 0, F(61, 225), F(-2, 9), F(13, 900)],
         [2, 0, 0, F(-53, 6), F(704, 45), F(-107, 9), F(67, 90), 3],
         [F(-91, 108), 0, 0, F(23, 108), F(-976, 135), F(311, 54), F(-19, 60), F(17, 6), F(-1, 12)],
         [F(2383, 4100), 0, 0, F(-341, 164), F(4496, 1025), F(-301, 82), F(2133, 4100), F(45, 82), F(45, 164),
          F(18, 41)],
         [F(3, 205), 0, 0, 0, 0, F(-6, 41), F(-3, 205), F(-3, 41), F(3, 41), F(6, 41), 0],
         [F(-1777, 4100), 0, 0, F(-341, 164), F(4496, 1025), F(-289, 82), F(2193, 4100), F(51, 82), F(33, 164),
          F(12, 41), 0, 1]]
    b = [0, 0, 0, 0, 0, F(34, 105), F(9, 35), F(9, 35), F(9, 280), F(9, 280), 0, F(41, 840), F(41, 840)]
    d = [F(41, 840), 0, 0, 0, 0, F(34, 105), F(9, 35), F(9, 35), F(9, 280), F(9, 280), F(41, 840), 0, 0]
    return "ARKODE_FEHLBERG_13_7_8", c, A, b, d, 8, 7


TABLES = {
    20: ("ARKODE_HEUN_EULER_2_1_2", [0, 1], [[1]], [F(1, 2), F(1, 2)], [1, 0], 2, 1),
    30: ("ARKODE_BOGACKI_SHAMPINE_4_2_3", [0, F(1, 2), F(3, 4), 1],
         [[F(1, 2)], [0, F(3, 4)], [F(2, 9), F(1, 3), F(4, 9)]],
         [F(2, 9), F(1, 3), F(4, 9), 0], [F(7, 24), F(1, 4), F(1, 3), F(1, 8)], 3, 2),
    31: _ark324(),
    40: ("ARKODE_ZONNEVELD_5_3_4", [0, F(1, 2), F(1, 2), 1, F(3, 4)],
         [[F(1, 2)], [0, F(1, 2)], [0, 0, 1], [F(5, 32), F(7, 32), F(13, 32), F(-1, 32)]],
         [F(1, 6), F(1, 3), F(1, 3), F(1, 6), 0], [F(-1, 2), F(7, 3), F(7, 3), F(13, 6), F(-16, 3)], 4, 3),
    41: _ark436(),
    42: ("ARKODE_SAYFY_ABURUB_6_3_4", [0, F(1, 2), 1, 1, F(1, 2), 1],
         [[F(1, 2)], [-1, 2], [F(1, 6), F(2, 3), F(1, 6)], [F(137, 1000), F(226, 1000), F(137, 1000)],
          [F(452, 1000), F(-904, 1000), F(-548, 1000), 0, 2]],
         [F(1, 6), F(1, 3), F(1, 12), 0, F(1, 3), F(1, 12)], [F(1, 6), F(2, 3), F(1, 6), 0, 0, 0], 4, 3),
    50: ("ARKODE_CASH_KARP_6_4_5", [0, F(1, 5), F(3, 10), F(3, 5), 1, F(7, 8)],
         [[F(1, 5)], [F(3, 40), F(9, 40)], [F(3, 10), F(-9, 10), F(6, 5)],
          [F(-11, 54), F(5, 2), F(-70, 27), F(35, 27)],
          [F(1631, 55296), F(175, 512), F(575, 13824), F(44275, 110592), F(253, 4096)]],
         [F(37, 378), 0, F(250, 621), F(125, 594), 0, F(512, 1771)],
         [F(2825, 27648), 0, F(18575, 48384), F(13525, 55296), F(277, 14336), F(1, 4)], 5, 4),
    51: ("ARKODE_FEHLBERG_6_4_5", [0, F(1, 4), F(3, 8), F(12, 13), 1, F(1, 2)],
         [[F(1, 4)], [F(3, 32), F(9, 32)], [F(1932, 2197), F(-7200, 2197), F(7296, 2197)],
          [F(439, 216), -8, F(3680, 513), F(-845, 4104)],
          [F(-8, 27), 2, F(-3544, 2565), F(1859, 4104), F(-11, 40)]],
         [F(16, 135), 0, F(6656, 12825), F(28561, 56430), F(-9, 50), F(2, 55)],
         [F(25, 216), 0, F(1408, 2565), F(2197, 4104), F(-1, 5), 0], 5, 4),
    52: ("ARKODE_DORMAND_PRINCE_7_4_5", [0, F(1, 5), F(3, 10), F(4, 5), F(8, 9), 1, 1],
         [[F(1, 5)], [F(3, 40), F(9, 40)], [F(44, 45), F(-56, 15), F(32, 9)],
          [F(19372, 6561), F(-25360, 2187), F(64448, 6561), F(-212, 729)],
          [F(9017, 3168), F(-355, 33), F(46732, 5247), F(49, 176), F(-5103, 18656)],
          [F(35, 384), 0, F(500, 1113), F(125, 192), F(-2187, 6784), F(11, 84)]],
         [F(35, 384), 0, F(500, 1113), F(125, 192), F(-2187, 6784), F(11, 84), 0],
         [F(5179, 57600), 0, F(7571, 16695), F(393, 640), F(-92097, 339200), F(187, 2100), F(1, 40)], 5, 4),
    53: _ark548(),
    60: ("ARKODE_VERNER_8_5_6", [0, F(1, 6), F(4, 15), F(2, 3), F(5, 6), 1, F(1, 15), 1],
         [[F(1, 6)], [F(4, 75), F(16, 75)], [F(5, 6), F(-8, 3), F(5, 2)],
          [F(-165, 64), F(55, 6), F(-425, 64), F(85, 96)],
          [F(12, 5), -8, F(4015, 612), F(-11, 36), F(88, 255)],
          [F(-8263, 15000), F(124, 75), F(-643, 680), F(-81, 250), F(2484, 10625), 0],
          [F(3501, 1720), F(-300, 43), F(297275, 52632), F(-319, 2322), F(24068, 84065), 0, F(3850, 26703)]],
         [F(3, 40), 0, F(875, 2244), F(23, 72), F(264, 1955), 0, F(125, 11592), F(43, 616)],
         [F(13, 160), 0, F(2375, 5984), F(5, 16), F(12, 85), F(3, 44), 0, 0], 6, 5),
    80: _fehlberg78(),
}
MISSING = {43: "ARKODE_ARK437L2SA_ERK_7_3_4", 54: "ARKODE_ARK548L2SAb_ERK_8_4_5"}


def method(code: int):
    """The table for ode.rk_method = code; unknown codes take the default, as
    ReactorArkode.cpp:78-81 does; codes 43 and 54 raise (not provided)."""
    if code in MISSING:
        raise KeyError(f"rk_method {code} ({MISSING[code]}) is not provided")
    return TABLES.get(code, TABLES[DEFAULT_METHOD])


def dense(code: int):
    """(name, s, q, p, c[s], A[s][s] (full, zero above the diagonal), b[s], d[s]) as Fractions."""
    name, c, A, b, d, q, p = method(code)
    s = len(c)
    full = [[F(0)] * s]
    for row in A:
        full.append([F(x) for x in row] + [F(0)] * (s - len(row)))
    return name, s, q, p, [F(x) for x in c], full, [F(x) for x in b], [F(x) for x in d]


def fsal(code: int) -> bool:
    """First same as last: the last stage is taken at the new solution
    (last row of A = b, c_s = 1), so it is the next step's first stage."""
    _, s, _, _, c, A, b, _ = dense(code)
    return c[-1] == 1 and A[-1][:-1] == b[:-1] and b[-1] == 0


def _fmt(x: F) -> str:
    return repr(float(x))  # correctly rounded; repr round-trips


def emit_header(path: str) -> None:
    lines = ["// Generated by python -m pelephysics_amd.erk_tables: the explicit Runge-Kutta",
             "// tables of ReactorArkode (ode.rk_method codes, ReactorArkode.cpp:26-83) as",
             "// correctly rounded doubles.  s stages, order q (b), embedding order p (d).",
             "#pragma once", "", "namespace ppc {", "",
             f"constexpr int ERK_MAX_STAGES = {MAX_STAGES};", "",
             "struct ErkTableHost {", "  int code;", "  const char* name;", "  int s, q, p, fsal;",
             f"  double c[{MAX_STAGES}];", f"  double A[{MAX_STAGES * MAX_STAGES}];  // row-major s x s",
             f"  double b[{MAX_STAGES}];", f"  double d[{MAX_STAGES}];", "};", "",
             "inline constexpr ErkTableHost ERK_TABLES[] = {"]
    for code in sorted(TABLES):
        name, s, q, p, c, A, b, d = dense(code)
        flatA = [A[i][j] for i in range(s) for j in range(s)]
        lines.append(f"    {{{code}, \"{name}\", {s}, {q}, {p}, {1 if fsal(code) else 0},")
        lines.append("     {" + ", ".join(_fmt(x) for x in c) + "},")
        lines.append("     {" + ", ".join(_fmt(x) for x in flatA) + "},")
        lines.append("     {" + ", ".join(_fmt(x) for x in b) + "},")
        lines.append("     {" + ", ".join(_fmt(x) for x in d) + "}},")
    lines += ["};", "", f"constexpr int ERK_NUM_TABLES = {len(TABLES)};", f"constexpr int ERK_DEFAULT = {DEFAULT_METHOD};",
              "", "}  // namespace ppc", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


if __name__ == "__main__":
    emit_header(os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "erk_tables.h"))

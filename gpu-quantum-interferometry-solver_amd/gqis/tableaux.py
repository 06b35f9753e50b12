"""Butcher tableaux of the solvers, emitted into the kernels as compile-time constants.

Sources (published methods; the reference lists the same values for its fixed-step bodies at
``gqis/cuda_solvers.py:705-721, 910-969``):
* Tsit5: Ch. Tsitouras, Comput. Math. Appl. 62 (2011) 770-775, with its embedded estimate.
* DP5: J. R. Dormand, P. J. Prince, J. Comput. Appl. Math. 6 (1980) 19-26.
* DOP853: E. Hairer, S. P. Norsett, G. Wanner, Solving ODEs I (2nd ed.), the DOP853 code:
  main formula plus the 5th- and 3rd-order error estimators (``E5``, ``E3``).
* Alshina6: the 7-stage sixth-order formula of E. A. Alshina, E. M. Zaks, N. N. Kalitkin (2008).
* LSRK4: M. H. Carpenter, C. A. Kennedy, NASA TM-109112 (1994), 2N-storage (5,4) scheme.
* Anas5(w): the frequency-fitted 6-stage formula (Anastassi & Simos) whose a43 coefficient is
  computed from w*dt on the host.
"""

TSIT5 = {
    "c": (0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0),
    "a": (
        (),
        (0.161,),
        (-0.008480655492356989, 0.335480655492357),
        (2.8971530571054935, -6.359448489975075, 4.3622954328695815),
        (5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525),
        (5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
         -0.028269050394068383),
    ),
    "b": (0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742,
          -3.290069515436081, 2.324710524099774),
    # b - b_hat; the 7th entry weights the FSAL derivative f(y_new)
    "e": (-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
          -0.1447110071732629, 0.5823571654525552, -0.45808210592918697,
          0.015151515151515152),
}

DP5 = {
    "c": (0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0),
    "a": (
        (),
        (1 / 5,),
        (3 / 40, 9 / 40),
        (44 / 45, -56 / 15, 32 / 9),
        (19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729),
        (9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656),
    ),
    "b": (35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84),
}

ALSHINA6 = {
    "c": (0.0, 4 / 7, 5 / 7, 6 / 7, 0.276393202250021, 0.7236067977499789, 1.0),
    "a": (
        (),
        (4 / 7,),
        (1.0267857142857142, -0.3125),
        (0.9349206349206349, 0.2777777777777778, -0.35555555555555557),
        (0.18002567144208434, 0.14737940683961614, 0.016070163293314288, -0.0670820393249937),
        (-0.07979765856603138, 0.025290591998992584, -0.3516326202266813, 0.3207294901687516,
         0.8090169943749475),
        (0.4988599356197352, -0.8633499941930429, 1.6778122846668349, -1.2682372542187894,
         -0.42705098312484235, 1.381966011250105),
    ),
    "b": (1 / 12, 0.0, 0.0, 0.0, 5 / 12, 5 / 12, 1 / 12),
}

DOP853 = {
    "c": (0.0, 0.526001519587677318785587544488e-01, 0.789002279381515978178381316732e-01,
          0.118350341907227396726757197510, 0.281649658092772603273242802490,
          0.333333333333333333333333333333, 0.25, 0.307692307692307692307692307692,
          0.651282051282051282051282051282, 0.6, 0.857142857142857142857142857142, 1.0),
    "a": (
        (),
        (5.26001519587677318785587544488e-2,),
        (1.97250569845378994544595329183e-2, 5.91751709536136983633785987549e-2),
        (2.95875854768068491816892993775e-2, 0.0, 8.87627564304205475450678981324e-2),
        (2.41365134159266685502369798665e-1, 0.0, -8.84549479328286085344864962717e-1,
         9.24834003261792003115737966543e-1),
        (3.7037037037037037037037037037e-2, 0.0, 0.0, 1.70828608729473871279604482173e-1,
         1.25467687566822425016691814123e-1),
        (3.7109375e-2, 0.0, 0.0, 1.70252211019544039314978060272e-1,
         6.02165389804559606850219397283e-2, -1.7578125e-2),
        (3.70920001185047927108779319836e-2, 0.0, 0.0, 1.70383925712239993810214054705e-1,
         1.07262030446373284651809199168e-1, -1.53194377486244017527936158236e-2,
         8.27378916381402288758473766002e-3),
        (6.24110958716075717114429577812e-1, 0.0, 0.0, -3.36089262944694129406857109825,
         -8.68219346841726006818189891453e-1, 2.75920996994467083049415600797e1,
         2.01540675504778934086186788979e1, -4.34898841810699588477366255144e1),
        (4.77662536438264365890433908527e-1, 0.0, 0.0, -2.48811461997166764192642586468,
         -5.90290826836842996371446475743e-1, 2.12300514481811942347288949897e1,
         1.52792336328824235832596922938e1, -3.32882109689848629194453265587e1,
         -2.03312017085086261358222928593e-2),
        (-9.3714243008598732571704021658e-1, 0.0, 0.0, 5.18637242884406370830023853209,
         1.09143734899672957818500254654, -8.14978701074692612513997267357,
         -1.85200656599969598641566180701e1, 2.27394870993505042818970056734e1,
         2.49360555267965238987089396762, -3.0467644718982195003823669022),
        (2.27331014751653820792359768449, 0.0, 0.0, -1.05344954667372501984066689879e1,
         -2.00087205822486249909675718444, -1.79589318631187989172765950534e1,
         2.79488845294199600508499808837e1, -2.85899827713502369474065508674,
         -8.87285693353062954433549289258, 1.23605671757943030647266201528e1,
         6.43392746015763530355970484046e-1),
    ),
    "b": (5.42937341165687622380535766363e-2, 0.0, 0.0, 0.0, 0.0,
          4.45031289275240888144113950566, 1.89151789931450038304281599044,
          -5.8012039600105847814672114227, 3.1116436695781989440891606237e-1,
          -1.52160949662516078556178806805e-1, 2.01365400804030348374776537501e-1,
          4.47106157277725905176885569043e-2),
    # 5th-order estimator (er of the DOP853 code)
    "e5": (0.1312004499419488073250102996e-1, 0.0, 0.0, 0.0, 0.0,
           -0.1225156446376204440720569753e+1, -0.4957589496572501915214079952,
           0.1664377182454986536961530415e+1, -0.3503288487499736816886487290,
           0.3341791187130174790297318841, 0.8192320648511571246570742613e-1,
           -0.2235530786388629525884427845e-1),
    # b - bhh (3rd-order estimator); bhh1 = 0.2440944881889764, bhh2 = 0.7338466882816118,
    # bhh3 = 0.02205882352941176 on stages 1, 9, 12
}
DOP853["e3"] = tuple(
    b - {0: 0.244094488188976377952755905512, 8: 0.733846688281611857341361741547,
         11: 0.220588235294117647058823529412e-1}.get(i, 0.0)
    for i, b in enumerate(DOP853["b"]))

LSRK4 = {
    "a": (0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
          -3550918686646 / 2091501179385, -1275806237668 / 842570457699),
    "b": (1432997174477 / 9575080441755, 5161836677717 / 13612068292357,
          1720146321549 / 2090206949498, 3134564353537 / 4481467310338,
          2277821191437 / 14882151754819),
    "c": (0.0, 1432997174477 / 9575080441755, 2526269341429 / 6820363962896,
          2006345519317 / 3224310063776, 2802321613138 / 2924317926251),
}

# Anas5(w): 6-stage fitted tableau.  Row 6 depends linearly on the fitted coefficient p
# (solver_param[0], evaluated on the host from w*dt): a6j = a0_6j + p * a1_6j.
ANAS5 = {
    "c": (0.0, 0.1, 1 / 3, 2 / 3, 0.9, 1.0),
    "a": (
        (),
        (1 / 10,),
        (-2 / 9, 5 / 9),
        (28 / 9, -40 / 9, 2.0),
        (-11277 / 8000, 171 / 80, -459 / 2000, 3213 / 8000),
        (-4.0, 5.0, 0.0, 0.0, 0.0),
    ),
    "a1": {5: (-0.595, 0.0, 1.89, -2.295, 1.0)},
    "b": (23 / 216, 0.0, 63 / 136, 9 / 56, 1000 / 3213, -1 / 24),
}


def _fmt(x):
    return repr(float(x))


def emit_arrays(tab, prefix="GQ"):
    """C++ constexpr arrays for a tableau dict (a is padded to a square lower triangle)."""
    n = len(tab["c"])
    lines = [f"#define {prefix}_NSTAGES {n}"]
    lines.append(f"static constexpr double {prefix}_C[{n}] = {{{', '.join(_fmt(v) for v in tab['c'])}}};")
    rows = []
    for s in range(n):
        row = list(tab["a"][s]) + [0.0] * (n - len(tab["a"][s]))
        rows.append("{" + ", ".join(_fmt(v) for v in row) + "}")
    lines.append(f"static constexpr double {prefix}_A[{n}][{n}] = {{{', '.join(rows)}}};")
    lines.append(f"static constexpr double {prefix}_B[{n}] = {{{', '.join(_fmt(v) for v in tab['b'])}}};")
    for key in ("e", "e5", "e3"):
        if key in tab:
            vals = tab[key]
            lines.append(f"static constexpr double {prefix}_{key.upper()}[{len(vals)}] = "
                         f"{{{', '.join(_fmt(v) for v in vals)}}};")
    return "\n".join(lines)


FIXED = {"tsit5": TSIT5, "dp5": DP5, "alshina6": ALSHINA6, "dop853": DOP853, "anas5": ANAS5}
AB5_START_C = (0.5, 1.0)   # RK4 start-up rows: midpoint, end


def _rows(cvals):
    """Distinct non-zero stage times of one step (first appearance), always ending with 1."""
    rows = []
    for c in cvals:
        if c != 0.0 and all(abs(c - r) > 5e-14 for r in rows):
            rows.append(c)
    if all(abs(1.0 - r) > 5e-14 for r in rows):
        rows.append(1.0)
    return rows


def _row_of(c, rows):
    if c == 0.0:
        return -1
    return min(range(len(rows)), key=lambda i: abs(rows[i] - c))


def emit_fixed(name):
    """Constants of the fixed-step template for solver ``name`` (scheme, tableau, time rows)."""
    lines = []
    if name in FIXED:
        tab = FIXED[name]
        n = len(tab["c"])
        rows = _rows(tab["c"])
        lines.append("#define GQ_SCHEME 0")
        lines.append(emit_arrays(tab))
        a1 = tab.get("a1", {})
        mat = []
        for s in range(n):
            row = list(a1.get(s, ())) + [0.0] * (n - len(a1.get(s, ())))
            mat.append("{" + ", ".join(_fmt(v) for v in row) + "}")
        lines.append(f"static constexpr double GQ_A1[{n}][{n}] = {{{', '.join(mat)}}};")
        row_of = [_row_of(c, rows) for c in tab["c"]]
    elif name in ("lserk4", "lserk4f"):
        n = len(LSRK4["c"])
        rows = _rows(LSRK4["c"])
        lines.append("#define GQ_SCHEME 1")
        lines.append(f"#define GQ_NSTAGES {n}")
        lines.append(f"static constexpr double GQ_A[{n}][1] = "
                     f"{{{', '.join('{' + _fmt(v) + '}' for v in LSRK4['a'])}}};")
        lines.append(f"static constexpr double GQ_B[{n}] = {{{', '.join(_fmt(v) for v in LSRK4['b'])}}};")
        row_of = [_row_of(c, rows) for c in LSRK4["c"]]
    elif name == "ab5":
        rows = list(AB5_START_C)
        lines.append("#define GQ_SCHEME 2")
        lines.append("#define GQ_NSTAGES 1")
        row_of = [-1]
    else:
        raise ValueError(f"no fixed-step tableau for {name}")
    end_row = _row_of(1.0, rows)
    lines.append(f"#define GQ_NROWS {len(rows)}")
    lines.append(f"#define GQ_END_ROW {end_row}")
    lines.append(f"static constexpr double GQ_ROW_C[{len(rows)}] = {{{', '.join(_fmt(v) for v in rows)}}};")
    lines.append(f"static constexpr int GQ_ROW_OF[{len(row_of)}] = {{{', '.join(str(v) for v in row_of)}}};")
    return "\n".join(lines)


def rows_per_step(name):
    if name == "rk4":
        return 2
    if name in FIXED:
        return len(_rows(FIXED[name]["c"]))
    if name in ("lserk4", "lserk4f"):
        return len(_rows(LSRK4["c"]))
    if name == "ab5":
        return len(AB5_START_C)
    return 0

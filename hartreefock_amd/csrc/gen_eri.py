"""Generate straight-line HIP code for the Head-Gordon-Pople ERI scheme, one
struct per angular class (LA LB | LC LD) with LA>=LB, LC>=LD, LA+LB>=LC+LD.

The scheme is the one of the reference (HartreeFock/GaussianTwoElectrons.cpp):

  vertical recursion on centre A   [e+1_i|0]^(m) = PA_i [e]^(m) + WP_i [e]^(m+1)
                                          + e_i/(2 zeta) ([e-1_i]^(m) - rho/zeta [e-1_i]^(m+1))   (:98-133)
  electron transfer A -> C         (e, f+1_i) = delta_i (e,f) - zeta/eta (e+1_i, f)
                                          + e_i/(2 eta) (e-1_i, f) + f_i/(2 eta) (e, f-1_i)       (:137-192)
  contraction over primitives, then
  horizontal recursion on the bra  (a, b+1_i| = (a+1_i, b| + AB_i (a, b|                           (:198-273)
  horizontal recursion on the ket  |c, d+1_i) = |c+1_i, d) + CD_i |c, d)                           (:278-360)

with the recursion direction i chosen as the largest Cartesian component of the
target (ties x > y > z), as there.  Unlike the reference, which fills whole
index ranges, only the intermediates the final block depends on are emitted.

Transfer convention.  The reference fills bra rows e at ket level |f| only for
|e| >= |f|-1 (:152); rows below stay zero.  For (dd|dd) the rows |e|=2 at |f|=4
are needed, so its two-centre-bra (dd|dd) integrals differ from the exact ones.
`exact=False` reproduces the reference (those entries are the constant 0);
`exact=True` computes them.

Each generated struct provides
  prim(F, PA, WP, inv2z, rz, dl, zoe, inv2e, acc)   one primitive quartet, adds into acc
  hrr(acc, AB, CD, out)                              contracted block -> (ab|cd) with the
                                                     per-component normalisation applied
and the exact operation counts PRIM_FLOPS / HRR_FLOPS (one add or one multiply
= 1 flop; a fused multiply-add counts 2).

Usage: python gen_eri.py OUT_DIR [LMAX]

Writes OUT_DIR/eri_classes.inc (the structs), OUT_DIR/k_<class>_<mode>_<variant>.hip (one
kernel instance each), OUT_DIR/d_<class>.hip (the per-class launcher) and OUT_DIR/dispatch.inc
(the host-side class table with the flop counts).
"""
import math
import sys


def comps(L):
    """Cartesian components of shell L in canonical order (xx,xy,xz,yy,yz,zz ...),
    reference HartreeFock/QuantumNumbers.h:60-80."""
    out = []
    for l in range(L, -1, -1):
        for m in range(L - l, -1, -1):
            out.append((l, m, L - l - m))
    return out


def direction(c):
    mx = max(c)
    return 0 if c[0] == mx else (1 if c[1] == mx else 2)


def add(c, i, k=1):
    t = list(c)
    t[i] += k
    return tuple(t)


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def comp_norm(c):
    return 1.0 / math.sqrt(dfact(2 * c[0] - 1) * dfact(2 * c[1] - 1) * dfact(2 * c[2] - 1))


def cname(c):
    return "%d%d%d" % c


XYZ = "xyz"


class ClassGen:
    def __init__(self, la, lb, lc, ld, exact):
        self.la, self.lb, self.lc, self.ld = la, lb, lc, ld
        self.exact = exact
        self.L = la + lb + lc + ld
        self.E = [c for L in range(la, la + lb + 1) for c in comps(L)]
        self.F = [c for L in range(lc, lc + ld + 1) for c in comps(L)]
        self.eidx = {c: i for i, c in enumerate(self.E)}
        self.fidx = {c: i for i, c in enumerate(self.F)}
        self.flops_prim = 0
        self.flops_hrr = 0

    def needs_zero_rows(self):
        """True when some intermediate the final block depends on lies in the rows the
        reference leaves at zero (|e| < |f| - 1)."""
        if self.exact:
            return False
        seen = set()
        stack = [(e, f) for e in self.E for f in self.F]
        while stack:
            e, f = stack.pop()
            if (e, f) in seen or sum(f) == 0:
                continue
            seen.add((e, f))
            if sum(e) < sum(f) - 1:
                return True
            i = direction(f)
            f1 = add(f, i, -1)
            stack.append((e, f1))
            stack.append((add(e, i), f1))
            if e[i] > 0:
                stack.append((add(e, i, -1), f1))
            if f1[i] > 0:
                stack.append((e, add(f1, i, -1)))
        return False

    # ---------------- primitive part: VRR + ETR + accumulate ----------------
    def etr_zero(self, e, f):
        return (not self.exact) and sum(f) > 0 and sum(e) < sum(f) - 1

    def gen_prim(self):
        lines = []
        vrr_done = {}
        etr_done = {}
        fl = [0]

        def vrr(e, m):
            key = (e, m)
            if key in vrr_done:
                return vrr_done[key]
            if sum(e) == 0:
                vrr_done[key] = "F[%d]" % m
                return vrr_done[key]
            i = direction(e)
            e1 = add(e, i, -1)
            a = vrr(e1, m)
            b = vrr(e1, m + 1)
            expr = "PA%s * %s + WP%s * %s" % (XYZ[i], a, XYZ[i], b)
            fl[0] += 3
            if e1[i] > 0:
                e2 = add(e1, i, -1)
                c = vrr(e2, m)
                d = vrr(e2, m + 1)
                expr += " + %s(%s - rz * %s)" % (ifac(e1[i], "inv2z"), c, d)
                fl[0] += 5
            name = "v%s_%d" % (cname(e), m)
            lines.append("        const double %s = %s;" % (name, expr))
            vrr_done[key] = name
            return name

        def etr(e, f):
            key = (e, f)
            if key in etr_done:
                return etr_done[key]
            if sum(f) == 0:
                etr_done[key] = vrr(e, 0)
                return etr_done[key]
            if self.etr_zero(e, f):
                etr_done[key] = None
                return None
            i = direction(f)
            f1 = add(f, i, -1)
            terms = []
            t = etr(e, f1)
            if t is not None:
                terms.append("dl%s * %s" % (XYZ[i], t))
            t = etr(add(e, i), f1)
            if t is not None:
                terms.append("- zoe * %s" % t)
            if e[i] > 0:
                t = etr(add(e, i, -1), f1)
                if t is not None:
                    terms.append("+ %s%s" % (ifac(e[i], "inv2e"), t))
            if f1[i] > 0:
                t = etr(e, add(f1, i, -1))
                if t is not None:
                    terms.append("+ %s%s" % (ifac(f1[i], "inv2e"), t))
            if not terms:
                etr_done[key] = None
                return None
            expr = " ".join(terms)
            if expr.startswith("- "):
                expr = "-" + expr[2:]
            elif expr.startswith("+ "):
                expr = expr[2:]
            fl[0] += 2 * len(terms) - 1 + sum(1 for x in terms if "inv2e" in x)
            name = "t%s_%s" % (cname(e), cname(f))
            lines.append("        const double %s = %s;" % (name, expr))
            etr_done[key] = name
            return name

        for f in self.F:
            for e in self.E:
                v = etr(e, f)
                if v is not None:
                    lines.append("        acc[%d] += %s;" % (self.eidx[e] * len(self.F) + self.fidx[f], v))
                    fl[0] += 1
        self.flops_prim = fl[0]
        return lines

    # ---------------- contracted part: HRR bra, HRR ket, normalisation ----------------
    def gen_hrr(self):
        lines = []
        NF = len(self.F)
        A, B, C, D = comps(self.la), comps(self.lb), comps(self.lc), comps(self.ld)
        fl = 0
        # bra: (a, b | f) for every f in F
        bra = {}

        def hb(a, b, f):
            nonlocal fl
            key = (a, b, f)
            if key in bra:
                return bra[key]
            if sum(b) == 0:
                bra[key] = "acc[%d]" % (self.eidx[a] * NF + self.fidx[f])
                return bra[key]
            i = direction(b)
            b1 = add(b, i, -1)
            x = hb(add(a, i), b1, f)
            y = hb(a, b1, f)
            name = "h%s_%s_%s" % (cname(a), cname(b), cname(f))
            lines.append("        const double %s = %s + AB%s * %s;" % (name, x, XYZ[i], y))
            fl += 2
            bra[key] = name
            return name

        ket = {}

        def hk(a, b, c, d):
            nonlocal fl
            key = (a, b, c, d)
            if key in ket:
                return ket[key]
            if sum(d) == 0:
                ket[key] = hb(a, b, c)
                return ket[key]
            i = direction(d)
            d1 = add(d, i, -1)
            x = hk(a, b, add(c, i), d1)
            y = hk(a, b, c, d1)
            name = "k%s_%s_%s_%s" % (cname(a), cname(b), cname(c), cname(d))
            lines.append("        const double %s = %s + CD%s * %s;" % (name, x, XYZ[i], y))
            fl += 2
            ket[key] = name
            return name

        nb, nc, nd = len(B), len(C), len(D)
        for ia, a in enumerate(A):
            for ib, b in enumerate(B):
                for ic, c in enumerate(C):
                    for id_, d in enumerate(D):
                        v = hk(a, b, c, d)
                        s = comp_norm(a) * comp_norm(b) * comp_norm(c) * comp_norm(d)
                        idx = ((ia * nb + ib) * nc + ic) * nd + id_
                        if s == 1.0:
                            lines.append("        out[%d] = %s;" % (idx, v))
                        else:
                            lines.append("        out[%d] = R(%.17g) * %s;" % (idx, s, v))
                            fl += 1
        self.flops_hrr = fl
        return lines

    def emit(self, truncated=False):
        name = "EriClass<%d, %d, %d, %d, %d>" % (self.la, self.lb, self.lc, self.ld, 1 if self.exact else 0)
        prim = self.gen_prim()
        hrr = self.gen_hrr()
        na, nb, nc, nd = (len(comps(x)) for x in (self.la, self.lb, self.lc, self.ld))
        out = []
        out.append("template <> struct %s {" % name)
        out.append("    static constexpr int L = %d, NA = %d, NB = %d, NC = %d, ND = %d;" % (self.L, na, nb, nc, nd))
        out.append("    static constexpr int NE = %d, NF = %d, NACC = %d, NOUT = %d;" %
                   (len(self.E), len(self.F), len(self.E) * len(self.F), na * nb * nc * nd))
        out.append("    static constexpr int PRIM_FLOPS = %d, HRR_FLOPS = %d;" % (self.flops_prim, self.flops_hrr))
        # TRUNCATED: the reference convention drops transfer rows for this class, so the value depends
        # on which pair is the bra; kernels then pick the bra per component as the reference does.
        out.append("    static constexpr bool TRUNCATED = %s;" % ("true" if truncated else "false"))
        out.append("    template <typename R>")
        out.append("    __host__ __device__ __forceinline__ static void prim(const R* __restrict__ F, R PAx, R PAy, R PAz, "
                   "R WPx, R WPy, R WPz, R inv2z, R rz, R dlx, R dly, R dlz, R zoe, R inv2e, R* __restrict__ acc) {")
        for l in prim:
            out.append(l.replace("const double ", "const R "))
        out.append("        (void)PAx; (void)PAy; (void)PAz; (void)WPx; (void)WPy; (void)WPz; (void)inv2z; (void)rz;")
        out.append("        (void)dlx; (void)dly; (void)dlz; (void)zoe; (void)inv2e;")
        out.append("    }")
        out.append("    template <typename R>")
        out.append("    __host__ __device__ __forceinline__ static void hrr(const R* __restrict__ acc, R ABx, R ABy, R ABz, "
                   "R CDx, R CDy, R CDz, R* __restrict__ out) {")
        for l in hrr:
            out.append(l.replace("const double ", "const R "))
        out.append("        (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;")
        out.append("    }")
        out.append("};")
        return "\n".join(out)


def classes(lmax):
    pairs = [(a, b) for a in range(lmax + 1) for b in range(a + 1)]
    pairs.sort(key=lambda p: (p[0] + p[1], p[0]))
    out = []
    for i, (la, lb) in enumerate(pairs):
        for j in range(i + 1):
            lc, ld = pairs[j]
            out.append((la, lb, lc, ld))
    return pairs, out


def main():
    outdir = sys.argv[1]
    lmax = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    import os
    os.makedirs(outdir, exist_ok=True)
    pairs, cls = classes(lmax)
    parts = ["// Generated by gen_eri.py - do not edit.", "#pragma once", ""]
    parts.append("template <int LA, int LB, int LC, int LD, int EXACT> struct EriClass;")
    parts.append("")
    info = []
    for (la, lb, lc, ld) in cls:
        ref = ClassGen(la, lb, lc, ld, False)
        ex = ClassGen(la, lb, lc, ld, True)
        ex_text = ex.emit()
        differs = any(ref.etr_zero(e, f) for e in ref.E for f in ref.F) or ref.needs_zero_rows()
        parts.append(ref.emit(truncated=differs))
        parts.append(ex_text)
        parts.append("")
        info.append((la, lb, lc, ld, differs, ref.flops_prim, ex.flops_prim, ref.flops_hrr))
    write_if_changed(os.path.join(outdir, "eri_classes.inc"), "\n".join(parts))
    decl = []
    table = []
    coop_flops = {}
    for (la, lb, lc, ld, differs, fr, fe, fh) in info:
        tag = "%d%d%d%d" % (la, lb, lc, ld)
        diag = (la, lb) == (lc, ld)
        sig = "(const TaskList& tl, const ShellsDev& sh, const PrimPair* pp, const double* boys, " \
              "const SinkArgs& s, hipStream_t st)"
        # one translation unit per (mode, variant) so the large classes compile in parallel
        variants = []
        coop = la + lb + lc + ld >= COOP_MIN_L
        if coop:
            for ex in ((0, 1) if differs else (0,)):
                cg = CoopGen(la, lb, lc, ld, bool(ex)).build()
                cg.truncated = differs and not ex
                write_if_changed(os.path.join(outdir, "coop_%s_%d.inc" % (tag, ex)), cg.emit("%s_%d" % (tag, ex)))
                coop_flops[(la, lb, lc, ld, ex)] = (cg.flops_prim, cg.flops_hrr)
        modes = ("MODE_JK", "MODE_PACKED") + (("MODE_DIAG",) if diag else ()) + (() if coop else ("MODE_JK32",))
        for mode in modes:
            for ex in ((0, 1) if differs else (0,)):
                fn = "run_%s_%s_%d" % (tag, mode.split("_")[1].lower(), ex)
                variants.append((mode, ex, fn))
                if coop:
                    inc = '#include "coop_%s_%d.inc"' % (tag, ex)
                    call = "launch_coop<CoopClass<%d, %d, %d, %d, %d>, %s>" % (la, lb, lc, ld, ex, mode)
                else:
                    inc = '#include "../eri_kernel.hpp"'
                    call = "launch_eri<%d, %d, %d, %d, %d, %s>" % (la, lb, lc, ld, ex, mode)
                body = ["// Generated by gen_eri.py - do not edit.", inc,
                        "namespace hfgpu {",
                        "hipError_t %s%s {" % (fn, sig),
                        "    return %s(tl, sh, pp, boys, s, st);" % call,
                        "}", "}  // namespace hfgpu"]
                write_if_changed(os.path.join(outdir, "k_%s.hip" % fn[4:]), "\n".join(body) + "\n")
        body = ["// Generated by gen_eri.py - do not edit.", '#include "../hf_device.h"', "namespace hfgpu {"]
        for (_, _, fn) in variants:
            body.append("hipError_t %s%s;" % (fn, sig))
        body.append("hipError_t launch_class_%s(int mode, int exact, const TaskList& tl, const ShellsDev& sh, "
                    "const PrimPair* pp, const double* boys, const SinkArgs& s, hipStream_t st) {" % tag)
        if coop:
            body.append("    if (mode == MODE_JK32) mode = MODE_JK;   // no FP32 variant of the LDS kernels")
        for (mode, ex, fn) in variants:
            cond = "mode == %s" % mode
            if differs:
                cond += " && %sexact" % ("" if ex else "!")
            body.append("    if (%s) return %s(tl, sh, pp, boys, s, st);" % (cond, fn))
        body.append("    (void)exact;")
        body.append("    return hipErrorInvalidValue;")
        body.append("}")
        body.append("}  // namespace hfgpu")
        write_if_changed(os.path.join(outdir, "d_%s.hip" % tag), "\n".join(body) + "\n")
        decl.append("hipError_t launch_class_%s(int, int, const TaskList&, const ShellsDev&, const PrimPair*, "
                    "const double*, const SinkArgs&, hipStream_t);" % tag)
        table.append("    {%d, %d, %d, %d, %d, %d, %d, %d, launch_class_%s}," % (la, lb, lc, ld, 1 if differs else 0,
                                                                             fr, fe, fh, tag))
    disp = ["// Generated by gen_eri.py - do not edit.", "#pragma once", '#include "../hf_device.h"',
            "namespace hfgpu {"]
    disp += decl
    disp.append("typedef hipError_t (*LaunchFn)(int, int, const TaskList&, const ShellsDev&, const PrimPair*, "
                "const double*, const SinkArgs&, hipStream_t);")
    disp.append("struct ClassInfo { int la, lb, lc, ld, exact_differs, prim_flops_ref, prim_flops_exact, hrr_flops; "
                "LaunchFn fn; };")
    disp.append("constexpr int kLmax = %d;" % lmax)
    disp.append("constexpr int kCoopMinL = %d;   // classes with LA+LB+LC+LD >= this use eri_coop_kernel" % COOP_MIN_L)
    disp.append("constexpr int kNumPairClasses = %d;" % len(pairs))
    disp.append("static const int kPairClassL[%d][2] = {%s};" % (len(pairs), ", ".join("{%d, %d}" % p for p in pairs)))
    disp.append("static const ClassInfo kClasses[] = {")
    disp += table
    disp.append("};")
    disp.append("constexpr int kNumClasses = %d;" % len(info))
    disp.append("}  // namespace hfgpu")
    write_if_changed(os.path.join(outdir, "dispatch.inc"), "\n".join(disp) + "\n")
    # host-side evaluation switch (CPU check of the generated code; used by tests only)
    hv = ["// Generated by gen_eri.py - do not edit.", "#pragma once", '#include "../eri_kernel.hpp"',
          "namespace hfgpu {",
          "inline int eval_quartet_host(int la, int lb, int lc, int ld, int exact, const double* A, const double* B, "
          "const double* C, const double* D, const PrimPair* bra, int npb, const PrimPair* ket, int npk, "
          "const double* boys, double* out, int sA, int sC, int a0, int b0, int c0, int d0) {"]
    for (la, lb, lc, ld, differs, fr, fe, fh) in info:
        for ex in ((0, 1) if differs else (0,)):
            cond = "la == %d && lb == %d && lc == %d && ld == %d" % (la, lb, lc, ld)
            if differs:
                cond += " && %sexact" % ("" if ex else "!")
            cls = "EriClass<%d, %d, %d, %d, %d>" % (la, lb, lc, ld, ex)
            if differs and not ex:   # reference convention: per-component bra choice, as the kernels do
                hv.append("    if (%s) { eval_quartet_reference_order<%s>(sA, sC, A, B, C, D, bra, npb, ket, npk, "
                          "a0, b0, c0, d0, boys, out); return 0; }" % (cond, cls))
            else:
                hv.append("    if (%s) { eval_quartet<%s>(A, B, C, D, bra, npb, ket, npk, boys, out); return 0; }"
                          % (cond, cls))
    hv.append("    return -1;")
    hv.append("}")
    hv.append("inline int eval_quartet_host_f32(int la, int lb, int lc, int ld, const double* A, const double* B, "
              "const double* C, const double* D, const PrimPair* bra, int npb, const PrimPair* ket, int npk, "
              "const float* boys32, float* out) {")
    for (la, lb, lc, ld, differs, fr, fe, fh) in info:
        if la + lb + lc + ld >= COOP_MIN_L:
            continue
        cond = "la == %d && lb == %d && lc == %d && ld == %d" % (la, lb, lc, ld)
        hv.append("    if (%s) { eval_quartet<EriClass<%d, %d, %d, %d, 0>, float>(A, B, C, D, bra, npb, ket, npk, "
                  "boys32, out); return 0; }" % (cond, la, lb, lc, ld))
    hv.append("    return -1;")
    hv.append("}")
    hv.append("}  // namespace hfgpu")
    write_if_changed(os.path.join(outdir, "host_eval.inc"), "\n".join(hv) + "\n")


def ifac(k, name):
    """`k * name * ` with an integer factor, so the product keeps the arithmetic type R."""
    return "%s * " % name if k == 1 else "%d * %s * " % (k, name)


def write_if_changed(path, text):
    """Keep timestamps (and so incremental builds) stable when nothing changed."""
    try:
        with open(path) as fh:
            if fh.read() == text:
                return
    except OSError:
        pass
    with open(path, "w") as fh:
        fh.write(text)



# ---------------------------------------------------------------------------------------
# Cooperative (one wavefront per quartet) form for the high-L classes.
#
# The same recursions as ClassGen, expressed as a dependency-levelled list of ops
#     val[dst] = sum_{r<4} coef[c_r] * val[s_r]
# over an LDS array `val`; the 64 lanes of a wave execute the ops of one level in parallel.
# Coefficient ids index a small per-primitive-quartet array (see COEF below).
# ---------------------------------------------------------------------------------------

COOP_MIN_L = 6      # classes with LA+LB+LC+LD >= this run one wavefront per quartet

COEF = {"zero": 0, "one": 1, "PA": 2, "WP": 5, "DL": 8, "mzoe": 11, "kinv2z": 12, "mrzkinv2z": 20,
        "kinv2e": 28, "AB": 36, "CD": 39}
NCOEF = 42


class CoopGen:
    def __init__(self, la, lb, lc, ld, exact):
        self.g = ClassGen(la, lb, lc, ld, exact)
        self.la, self.lb, self.lc, self.ld = la, lb, lc, ld
        self.exact = exact

    def build(self):
        g = self.g
        L = g.L
        # ---- phase 1: per primitive quartet ----
        nodes = {}          # key -> (level, terms) ; terms = [(coef_id, src_key)]
        base = {}

        def vrr(e, m):
            key = ("v", e, m)
            if key in nodes or key in base:
                return key
            if sum(e) == 0:
                base[key] = m
                return key
            i = direction(e)
            e1 = add(e, i, -1)
            terms = [(COEF["PA"] + i, vrr(e1, m)), (COEF["WP"] + i, vrr(e1, m + 1))]
            if e1[i] > 0:
                e2 = add(e1, i, -1)
                terms.append((COEF["kinv2z"] + e1[i] - 1, vrr(e2, m)))
                terms.append((COEF["mrzkinv2z"] + e1[i] - 1, vrr(e2, m + 1)))
            nodes[key] = terms
            return key

        def etr(e, f):
            if sum(f) == 0:
                return vrr(e, 0)
            key = ("t", e, f)
            if key in nodes:
                return key
            if g.etr_zero(e, f):
                return None
            i = direction(f)
            f1 = add(f, i, -1)
            terms = []
            t = etr(e, f1)
            if t is not None:
                terms.append((COEF["DL"] + i, t))
            t = etr(add(e, i), f1)
            if t is not None:
                terms.append((COEF["mzoe"], t))
            if e[i] > 0:
                t = etr(add(e, i, -1), f1)
                if t is not None:
                    terms.append((COEF["kinv2e"] + e[i] - 1, t))
            if f1[i] > 0:
                t = etr(e, add(f1, i, -1))
                if t is not None:
                    terms.append((COEF["kinv2e"] + f1[i] - 1, t))
            if not terms:
                return None
            nodes[key] = terms
            return key

        NF = len(g.F)
        acc_src = {}
        for f in g.F:
            for e in g.E:
                k = etr(e, f)
                if k is not None:
                    acc_src[g.eidx[e] * NF + g.fidx[f]] = k
        level = {}

        def lev(k):
            if k in base:
                return 0
            if k in level:
                return level[k]
            lv = 1 + max(lev(s) for _, s in nodes[k])
            level[k] = lv
            return lv

        for k in nodes:
            lev(k)
        # LDS slots: 0 = constant zero; [1, 1+NACC) the contracted block (persistent across
        # primitive quartets); after it the phase-1 region (F_m, VRR, transfer values) and, reusing
        # the same space once the primitive loop is over, the phase-2 region (HRR values).
        NFc = len(g.F)
        nacc = len(g.E) * NFc
        acc0 = 1
        slot = {}
        nxt = acc0 + nacc
        for k, m in sorted(base.items(), key=lambda kv: kv[1]):
            slot[k] = nxt
            nxt += 1
        self.F0 = slot[("v", (0, 0, 0), 0)]
        order = sorted(nodes, key=lambda k: (level[k], str(k)))
        for k in order:
            slot[k] = nxt
            nxt += 1
        phase1_end = nxt
        # ---- phase 2: HRR after contraction ----
        A, B, C, D = comps(self.la), comps(self.lb), comps(self.lc), comps(self.ld)
        h_nodes = {}
        h_base = {}

        def hb(a, b, f):
            key = ("hb", a, b, f)
            if key in h_nodes or key in h_base:
                return key
            if sum(b) == 0:
                h_base[key] = g.eidx[a] * NF + g.fidx[f]
                return key
            i = direction(b)
            b1 = add(b, i, -1)
            h_nodes[key] = [(COEF["one"], hb(add(a, i), b1, f)), (COEF["AB"] + i, hb(a, b1, f))]
            return key

        def hk(a, b, c, d):
            if sum(d) == 0:
                return hb(a, b, c)
            key = ("hk", a, b, c, d)
            if key in h_nodes:
                return key
            i = direction(d)
            d1 = add(d, i, -1)
            h_nodes[key] = [(COEF["one"], hk(a, b, add(c, i), d1)), (COEF["CD"] + i, hk(a, b, c, d1))]
            return key

        outs = []
        for a in A:
            for b in B:
                for c in C:
                    for d in D:
                        outs.append((hk(a, b, c, d), comp_norm(a) * comp_norm(b) * comp_norm(c) * comp_norm(d)))
        hlevel = {}

        def hlev(k):
            if k in h_base:
                return 0
            if k in hlevel:
                return hlevel[k]
            lv = 1 + max(hlev(s) for _, s in h_nodes[k])
            hlevel[k] = lv
            return lv

        for k in h_nodes:
            hlev(k)
        for k, ai in h_base.items():
            slot[k] = acc0 + ai
        h2 = acc0 + nacc
        horder = sorted(h_nodes, key=lambda k: (hlevel[k], str(k)))
        for k in horder:
            slot[k] = h2
            h2 += 1
        nslot = max(phase1_end, h2)
        # ---- op lists by level ----
        def ops_of(keys, nodes_, lvl):
            levels = {}
            for k in keys:
                levels.setdefault(lvl[k], []).append(k)
            out, bounds = [], [0]
            for lv in sorted(levels):
                for k in levels[lv]:
                    terms = nodes_[k] + [(COEF["zero"], None)] * (4 - len(nodes_[k]))
                    out.append((slot[k], [slot[s] if s is not None else 0 for _, s in terms],
                                [c for c, _ in terms]))
                bounds.append(len(out))
            return out, bounds

        p_ops, p_bounds = ops_of(order, nodes, level)
        # accumulation level: acc[i] += val[src]
        acc_ops = [(acc0 + i, [acc0 + i, slot[k], 0, 0], [COEF["one"], COEF["one"], 0, 0])
                   for i, k in sorted(acc_src.items())]
        p_ops += acc_ops
        p_bounds.append(len(p_ops))
        h_ops, h_bounds = ops_of(horder, h_nodes, hlevel)
        self.nslot = nslot
        self.acc0, self.nacc = acc0, nacc
        self.p_ops, self.p_bounds = p_ops, p_bounds
        self.h_ops, self.h_bounds = h_ops, h_bounds
        self.out_slot = [slot[k] for k, _ in outs]
        self.out_norm = [nv for _, nv in outs]
        self.flops_prim = sum(2 * sum(1 for c in cs if c != 0) - 1 for _, _, cs in p_ops)
        self.flops_hrr = sum(3 for _ in h_ops) + sum(1 for v in self.out_norm if v != 1.0)
        return self

    def emit(self, tag):
        def packed(ops):
            rows = []
            for d, s, c in ops:
                rows.append("{%d, {%d, %d, %d, %d}, {%d, %d, %d, %d}}" % (d, s[0], s[1], s[2], s[3], c[0], c[1], c[2], c[3]))
            return ",\n    ".join(rows) if rows else "{0, {0, 0, 0, 0}, {0, 0, 0, 0}}"
        la, lb, lc, ld = self.la, self.lb, self.lc, self.ld
        n = "CoopClass<%d, %d, %d, %d, %d>" % (la, lb, lc, ld, 1 if self.exact else 0)
        out = ["// Generated by gen_eri.py - do not edit.", "#pragma once", '#include "../coop_kernel.hpp"',
               "namespace hfgpu {"]
        out.append("static __device__ const CoopOp coop_prim_%s[] = {\n    %s};" % (tag, packed(self.p_ops)))
        out.append("static __device__ const CoopOp coop_hrr_%s[] = {\n    %s};" % (tag, packed(self.h_ops)))
        out.append("static __device__ const unsigned short coop_out_%s[] = {%s};" %
                   (tag, ", ".join(str(x) for x in self.out_slot)))
        out.append("static __device__ const double coop_norm_%s[] = {%s};" %
                   (tag, ", ".join("%.17g" % x for x in self.out_norm)))
        out.append("template <> struct %s {" % n)
        na, nb, nc, nd = (len(comps(x)) for x in (la, lb, lc, ld))
        out.append("    static constexpr int L = %d, NA = %d, NB = %d, NC = %d, ND = %d, NOUT = %d;" %
                   (la + lb + lc + ld, na, nb, nc, nd, na * nb * nc * nd))
        out.append("    static constexpr int NSLOT = %d, F0 = %d, ACC0 = %d, NACC = %d;" %
                   (self.nslot, self.F0, self.acc0, self.nacc))
        # TRUNCATED: reference transfer convention; the kernel evaluates both bra choices and picks
        # per component as the reference does
        out.append("    static constexpr bool TRUNCATED = %s;" % ("true" if getattr(self, "truncated", False) else "false"))
        out.append("    static constexpr int NPL = %d, NHL = %d;" % (len(self.p_bounds) - 1, len(self.h_bounds) - 1))
        out.append("    static constexpr int PL[%d] = {%s};" % (len(self.p_bounds), ", ".join(map(str, self.p_bounds))))
        out.append("    static constexpr int HL[%d] = {%s};" % (len(self.h_bounds), ", ".join(map(str, self.h_bounds))))
        out.append("    __device__ static const CoopOp* prim_ops() { return coop_prim_%s; }" % tag)
        out.append("    __device__ static const CoopOp* hrr_ops() { return coop_hrr_%s; }" % tag)
        out.append("    __device__ static const unsigned short* out_slot() { return coop_out_%s; }" % tag)
        out.append("    __device__ static const double* out_norm() { return coop_norm_%s; }" % tag)
        out.append("};")
        out.append("}  // namespace hfgpu")
        return "\n".join(out) + "\n"


if __name__ == "__main__":
    main()

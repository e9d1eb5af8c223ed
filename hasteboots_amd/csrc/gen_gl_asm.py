"""Generates hb_gl_asm.inc: hand-scheduled gfx950 VALU sequences for Goldilocks.

Run `python gen_gl_asm.py` in this directory after editing; a CPU test checks that the
committed hb_gl_asm.inc matches this generator's output.

Arithmetic (P = 2^64 - 2^32 + 1, EPS = 2^32 - 1, words are 64-bit pairs of 32-bit halves):
  montmul(y, w): y in [0, 2^64), w < P stored times 2^64 (Montgomery constant form):
    (lo, hi) = y * w from four 32x32 products chained through v_mad_u64_u32, then the
    Goldilocks Montgomery reduction
        a = lo + (lo << 32) mod 2^64 (carry e);  b = a - (a >> 32) - e;
        r = hi - b (borrow c);  r -= c ? EPS : 0
    which returns y * (w / 2^64) mod P in [0, P) whenever hi < P (hi < w < P here).
  CT butterfly (forward):  t = montmul(y, w);  x' = x + t;  y' = x - t.
    t < P, so each of the add and the subtract needs one conditional EPS correction;
    x' and y' are in [0, 2^64).
  GS butterfly (inverse), x and y canonical:  x' = canon(x + y);  y' = montmul(x - y mod P, w).
  MAC:  p' = p + montmul(x, k) with p in [0, 2^64); one EPS correction.
Conditional +-EPS uses v_mad_i64_i32(m, 0x55555555, v) with m in {3, -3, 0}
(3 * 0x55555555 = EPS), or v_mad_u64_u32(m, 1, v) with m in {EPS, 0}.

Scheduling: two independent chains are interleaved by a list scheduler. gfx950 needs
two wait states between a VALU write of an SGPR (carry out / compare) and a VALU read
of it as carry-in or select mask; the scheduler fills them with the other chain's work
and emits s_nop only when nothing is ready. Temporaries live in fixed VGPRs
v232..v255 (clobbered); inputs and outputs are compiler-allocated operands, and every
output is written after the last read of every input, so outputs may share registers
with inputs.
"""
from __future__ import annotations

import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "hb_gl_asm.inc")

TEMP_BASE = 232  # v232..v255: 12 per chain
# One zero-extension pair shared by both chains (one fewer mov per block, but the two
# chains' products then run one after the other); measured slower on MI355X, so off.
SHARED_Z = False
# Preferred issue slots between a v_mad_u64_u32 and an instruction reading its result
# (a soft scheduling constraint). On MI355X back-to-back dependent mads cost ~6.6 clocks
# each and a gap of three instructions restores the full rate (4.2), but with two waves
# per SIMD the other wave fills those stalls: a preference of 3 measured no faster than
# none, so it is off.
MAD_LATENCY = 0
# Form the first zero-extended word of a product, (A.hi, 0), with one 64-bit shift of A
# instead of a move into Z.lo: the shift also writes Z.hi = 0, so no block has to zero it
# (2 fewer instructions per block). Measured no faster on MI355X (4096 Goldilocks gates
# 211.6 against 211.2 ms), so off.
ZEXT_SHIFT = False


class Chain:
    """Instruction list of one butterfly/MAC chain with symbolic registers.

    Symbolic names: 'A','U','V','W','Z' are fixed pairs (".lo"/".hi" halves), 'T1','T2'
    fixed singles, 'S1','S2' SGPR-pair operands, '%name' compiler operands, 'D' dummy
    SGPR pair. Each instruction is (template, defs, uses).
    """

    def __init__(self, tag: str, base: int):
        self.tag = tag
        self.base = base
        self.ins: list[tuple[str, list[str], list[str]]] = []

    def reg(self, name: str) -> str:
        pairs = {"A": 0, "U": 2, "V": 4, "W": 6, "Z": 8}
        if name in ("T1", "T2"):
            return f"v{self.base + 10 + (0 if name == 'T1' else 1)}"
        m = re.fullmatch(r"([AUVWZ])(\.lo|\.hi)?", name)
        if m:
            # Z (the zero-extension pair, high half kept 0) is shared by all chains
            shared = SHARED_Z and m.group(1) == "Z"
            b = (TEMP_BASE if shared else self.base) + pairs[m.group(1)]
            if m.group(2) == ".lo":
                return f"v{b}"
            if m.group(2) == ".hi":
                return f"v{b + 1}"
            return f"v[{b}:{b + 1}]"
        raise KeyError(name)

    def add(self, text: str, defs: list[str], uses: list[str]):
        self.ins.append((text, defs, uses))


def is_sgpr(sym: str) -> bool:
    return sym.startswith("S") or sym == "D" or sym.startswith("%s")


def halves(sym: str) -> list[str]:
    """Register atoms of a symbol, for dependence tracking (pairs -> both halves)."""
    if re.fullmatch(r"[AUVWZ]", sym):
        return [sym + ".lo", sym + ".hi"]
    return [sym]


def zext_a(c: Chain):
    """Z <- (A.hi, 0): Z's low word for the next product's addend (see ZEXT_SHIFT)."""
    if ZEXT_SHIFT:
        c.add("v_lshrrev_b64 {Z}, 32, {A}", ["Z"], ["A"])
    else:
        c.add("v_mov_b32 {Z.lo}, {A.hi}", ["Z.lo"], ["A.hi"])


def montmul_w(c: Chain, w0: str, w1: str, dst: str):
    """dst <- montmul(y, w) with y in the fixed pair W (consumed): y0 * w0 and y0 * w1 are
    formed first, y1 * w1 last, so W is rewritten only after its last read."""
    c.add(f"v_mad_u64_u32 {{A}}, {{D}}, {{W.lo}}, {w0}, 0", ["A"], ["W.lo", w0])
    c.add(f"v_mad_u64_u32 {{V}}, {{D}}, {{W.lo}}, {w1}, 0", ["V"], ["W.lo", w1])
    zext_a(c)
    c.add(f"v_mad_u64_u32 {{U}}, {{D}}, {{W.hi}}, {w0}, {{Z}}", ["U"], ["W.hi", w0, "Z"])
    c.add("v_mad_u64_u32 {V}, {D}, {U.lo}, 1, {V}", ["V"], ["U.lo", "V"])
    c.add("v_mov_b32 {Z.lo}, {U.hi}", ["Z.lo"], ["U.hi"])
    c.add(f"v_mad_u64_u32 {{W}}, {{D}}, {{W.hi}}, {w1}, {{Z}}", ["W"], ["W.hi", w1, "Z"])
    c.add("v_mad_u64_u32 {W}, {D}, {V.hi}, 1, {W}", ["W"], ["V.hi", "W"])
    monty(c, "A.lo", "V.lo", "W.lo", "W.hi", dst)


def montmul(c: Chain, y0: str, y1: str, w0: str, w1: str, dst: str = "A"):
    """dst (default pair A) <- montmul(y, w); uses A, U, V, W, Z, T1, T2, S1, S2."""
    c.add(f"v_mad_u64_u32 {{A}}, {{D}}, {y0}, {w0}, 0", ["A"], [y0, w0])
    zext_a(c)
    c.add(f"v_mad_u64_u32 {{U}}, {{D}}, {y1}, {w0}, {{Z}}", ["U"], [y1, w0, "Z"])
    c.add("v_mov_b32 {Z.lo}, {U.lo}", ["Z.lo"], ["U.lo"])
    c.add(f"v_mad_u64_u32 {{V}}, {{D}}, {y0}, {w1}, {{Z}}", ["V"], [y0, w1, "Z"])
    c.add("v_mov_b32 {Z.lo}, {U.hi}", ["Z.lo"], ["U.hi"])
    c.add(f"v_mad_u64_u32 {{W}}, {{D}}, {y1}, {w1}, {{Z}}", ["W"], [y1, w1, "Z"])
    c.add("v_mad_u64_u32 {W}, {D}, {V.hi}, 1, {W}", ["W"], ["V.hi", "W"])
    # lo = (A.lo, V.lo), hi = (W.lo, W.hi)
    monty(c, "A.lo", "V.lo", "W.lo", "W.hi", dst)


def word(sym: str) -> str:
    """Asm text of a 32-bit word symbol: a fixed register, an operand or the constant 0."""
    if sym == "0":
        return "0"
    return sym if sym.startswith("%") else "{" + sym + "}"


def monty(c: Chain, l0: str, l1: str, h0: str, h1: str, dst: str):
    """dst <- (lo + hi * 2^64) * 2^-64 mod P, canonical (hi < P); words may be "0".
    a = lo + (lo << 32) (carry e); b = a - (a >> 32) - e; r = hi - b (borrow c);
    r -= c ? EPS : 0. Writes A, T1, T2, S1, S2 after reading its inputs."""
    w = word
    c.add(f"v_add_co_u32 {{T1}}, {{S1}}, {w(l1)}, {w(l0)}", ["T1", "S1"], [l1, l0])
    c.add(f"v_subb_co_u32 {{T2}}, {{S2}}, {w(l0)}, {{T1}}, {{S1}}", ["T2", "S2"], [l0, "T1", "S1"])
    c.add("v_subb_co_u32 {T1}, {S2}, {T1}, 0, {S2}", ["T1", "S2"], ["T1", "S2"])
    c.add(f"v_sub_co_u32 {{A.lo}}, {{S1}}, {w(h0)}, {{T2}}", ["A.lo", "S1"], [h0, "T2"])
    c.add(f"v_subb_co_u32 {{A.hi}}, {{S1}}, {w(h1)}, {{T1}}, {{S1}}", ["A.hi", "S1"], [h1, "T1", "S1"])
    c.add("v_cndmask_b32 {T2}, 0, -3, {S1}", ["T2"], ["S1"])
    c.add(f"v_mad_i64_i32 {{{dst}}}, {{D}}, {{T2}}, {{%k55}}, {{A}}", [dst], ["T2", "%k55", "A"])


def mulpow2(c: Chain, typ: str, y0: str, y1: str, dst: str):
    """dst <- y * 2^e mod P, canonical, y in [0, 2^64) given as words y0, y1; the shift
    amounts are the chain's immediate operands %s<tag>1, %s<tag>2.
      A  (e = k, 1 <= k < 32; s1 = k, s2 = 32 - k): y * 2^k = W0 + W1 * 2^32 + W2 * 2^64
         with W2 < 2^32, so t = (W0, W1) + W2 * EPS (carry c), q = t + EPS (carry c2),
         result q if c or c2 else t.
      BL (1 <= j < 32; s1 = j, s2 = 32 - j) and BH (32 < j < 64; s1 = j - 32, s2 = 64 - j):
         the Montgomery reduction of the 128-bit y * 2^j, i.e. y * 2^(j - 64)."""
    s1, s2 = "{%s" + c.tag + "1}", "{%s" + c.tag + "2}"
    y0t, y1t = word(y0), word(y1)
    if typ == "A":
        c.add(f"v_lshlrev_b32 {{A.lo}}, {s1}, {y0t}", ["A.lo"], [y0])
        c.add(f"v_lshrrev_b32 {{T1}}, {s2}, {y1t}", ["T1"], [y1])
        c.add(f"v_alignbit_b32 {{A.hi}}, {y1t}, {y0t}, {s2}", ["A.hi"], [y1, y0])
        c.add("v_mad_u64_u32 {A}, {S1}, {T1}, -1, {A}", ["A", "S1"], ["T1", "A"])
        c.add("v_mad_u64_u32 {U}, {S2}, -1, 1, {A}", ["U", "S2"], ["A"])
        c.add("s_or_b64 {S1}, {S1}, {S2}", ["S1"], ["S1", "S2"])
        c.add("v_cndmask_b32 {A.lo}, {A.lo}, {U.lo}, {S1}", ["A.lo"], ["A.lo", "U.lo", "S1"])
        c.add("v_cndmask_b32 {A.hi}, {A.hi}, {U.hi}, {S1}", ["A.hi"], ["A.hi", "U.hi", "S1"])
        if dst != "A":
            c.add(f"v_mov_b64 {{{dst}}}, {{A}}", [dst], ["A"])
    elif typ == "BL":
        c.add(f"v_lshlrev_b32 {{A.lo}}, {s1}, {y0t}", ["A.lo"], [y0])
        c.add(f"v_lshrrev_b32 {{W.lo}}, {s2}, {y1t}", ["W.lo"], [y1])
        c.add(f"v_alignbit_b32 {{V.lo}}, {y1t}, {y0t}, {s2}", ["V.lo"], [y1, y0])
        monty(c, "A.lo", "V.lo", "W.lo", "0", dst)
    elif typ == "BH":
        c.add(f"v_alignbit_b32 {{W.lo}}, {y1t}, {y0t}, {s2}", ["W.lo"], [y1, y0])
        c.add(f"v_lshrrev_b32 {{W.hi}}, {s2}, {y1t}", ["W.hi"], [y1])
        c.add(f"v_lshlrev_b32 {{V.lo}}, {s1}, {y0t}", ["V.lo"], [y0])
        monty(c, "0", "V.lo", "W.lo", "W.hi", dst)
    else:
        raise ValueError(typ)


def ct_p2_chain(tag: str, base: int, typ: str) -> Chain:
    """x' = x + t, y' = x - t with t = y * 2^e (mulpow2)."""
    c = Chain(tag, base)
    x0, x1, y0, y1 = f"%x{tag}0", f"%x{tag}1", f"%y{tag}0", f"%y{tag}1"
    mulpow2(c, typ, y0, y1, "A")
    ct_addsub(c, tag, x0, x1)
    return c


def ct_triv_chain(tag: str, base: int, typ: str) -> Chain:
    """Twiddle 1: x' = x + y, y' = x - y. TC: y canonical; TN: y in [0, 2^64), first made
    canonical (y + EPS carries exactly when y >= P, and then is y - P)."""
    c = Chain(tag, base)
    x0, x1, y0, y1 = f"%x{tag}0", f"%x{tag}1", f"%y{tag}0", f"%y{tag}1"
    if typ == "TC":
        ct_addsub(c, tag, x0, x1, y0, y1)
        return c
    c.add(f"v_add_co_u32 {{A.lo}}, {{S1}}, {y0}, -1", ["A.lo", "S1"], [y0])
    c.add(f"v_addc_co_u32 {{A.hi}}, {{S1}}, {y1}, 0, {{S1}}", ["A.hi", "S1"], [y1, "S1"])
    c.add(f"v_cndmask_b32 {{A.lo}}, {y0}, {{A.lo}}, {{S1}}", ["A.lo"], [y0, "A.lo", "S1"])
    c.add(f"v_cndmask_b32 {{A.hi}}, {y1}, {{A.hi}}, {{S1}}", ["A.hi"], [y1, "A.hi", "S1"])
    ct_addsub(c, tag, x0, x1)
    return c


def ct_any_chain(tag: str, base: int, typ: str) -> Chain:
    """CT butterfly whose twiddle is +-2^e (A, BL, BH) or 1 (TC, TN)."""
    return ct_triv_chain(tag, base, typ) if typ in ("TC", "TN") else ct_p2_chain(tag, base, typ)


def gs_p2_chain(tag: str, base: int, typ: str) -> Chain:
    """x, y canonical: x' = canon(x + y) (two halves out), y' = (x - y) * 2^e."""
    c = Chain(tag, base)
    x0, x1, y0, y1 = f"%x{tag}0", f"%x{tag}1", f"%y{tag}0", f"%y{tag}1"
    gs_diff(c, x0, x1, y0, y1)
    mulpow2(c, typ, "V.lo", "V.hi", f"%Y{tag}")
    gs_sum(c, tag, x0, x1, y0, y1)
    return c


def ct_chain(tag: str, base: int, wkind: str) -> Chain:
    """x' = x + montmul(y, w), y' = x - montmul(y, w)."""
    c = Chain(tag, base)
    x0, x1, y0, y1 = f"%x{tag}0", f"%x{tag}1", f"%y{tag}0", f"%y{tag}1"
    w0, w1 = f"%{wkind}w{tag}0", f"%{wkind}w{tag}1"
    montmul(c, y0, y1, w0, w1)
    ct_addsub(c, tag, x0, x1)
    return c


def ct_addsub(c: Chain, tag: str, x0: str, x1: str, a0: str = "A.lo", a1: str = "A.hi"):
    """Outputs %X = x + a, %Y = x - a for canonical a (default the pair A), one EPS
    correction each."""
    w0, w1 = word(a0), word(a1)
    c.add(f"v_add_co_u32 {{U.lo}}, {{S1}}, {x0}, {w0}", ["U.lo", "S1"], [x0, a0])
    c.add(f"v_sub_co_u32 {{V.lo}}, {{S2}}, {x0}, {w0}", ["V.lo", "S2"], [x0, a0])
    c.add(f"v_addc_co_u32 {{U.hi}}, {{S1}}, {x1}, {w1}, {{S1}}", ["U.hi", "S1"], [x1, a1, "S1"])
    c.add(f"v_subb_co_u32 {{V.hi}}, {{S2}}, {x1}, {w1}, {{S2}}", ["V.hi", "S2"], [x1, a1, "S2"])
    c.add("v_cndmask_b32 {T1}, 0, -1, {S1}", ["T1"], ["S1"])
    c.add("v_cndmask_b32 {T2}, 0, -3, {S2}", ["T2"], ["S2"])
    c.add(f"v_mad_u64_u32 {{%X{tag}}}, {{D}}, {{T1}}, 1, {{U}}", [f"%X{tag}"], ["T1", "U"])
    c.add(f"v_mad_i64_i32 {{%Y{tag}}}, {{D}}, {{T2}}, {{%k55}}, {{V}}", [f"%Y{tag}"],
          ["T2", "%k55", "V"])


def gs_chain(tag: str, base: int, wkind: str) -> Chain:
    """x, y canonical: x' = canon(x + y) (two halves out), y' = montmul(x - y, w)."""
    c = Chain(tag, base)
    x0, x1, y0, y1 = f"%x{tag}0", f"%x{tag}1", f"%y{tag}0", f"%y{tag}1"
    w0, w1 = f"%{wkind}w{tag}0", f"%{wkind}w{tag}1"
    # d = x - y mod P, canonical, into W; y' = montmul(d, w) straight into the output
    gs_diff(c, x0, x1, y0, y1, "W")
    montmul_w(c, w0, w1, f"%Y{tag}")
    gs_sum(c, tag, x0, x1, y0, y1)
    return c


def gs_diff(c: Chain, x0: str, x1: str, y0: str, y1: str, r: str = "V"):
    """Pair r <- x - y mod P, canonical for canonical x, y."""
    c.add(f"v_sub_co_u32 {{{r}.lo}}, {{S2}}, {x0}, {y0}", [f"{r}.lo", "S2"], [x0, y0])
    c.add(f"v_subb_co_u32 {{{r}.hi}}, {{S2}}, {x1}, {y1}, {{S2}}", [f"{r}.hi", "S2"],
          [x1, y1, "S2"])
    c.add("v_cndmask_b32 {T2}, 0, -3, {S2}", ["T2"], ["S2"])
    c.add(f"v_mad_i64_i32 {{{r}}}, {{D}}, {{T2}}, {{%k55}}, {{{r}}}", [r], ["T2", "%k55", r])


def gs_sum(c: Chain, tag: str, x0: str, x1: str, y0: str, y1: str):
    """Outputs %X<tag>lo/hi = canon(x + y): s = x + y (carry S1); q = s + EPS (carry S2);
    q if S1 | S2 else s."""
    c.add(f"v_add_co_u32 {{U.lo}}, {{S1}}, {x0}, {y0}", ["U.lo", "S1"], [x0, y0])
    c.add(f"v_addc_co_u32 {{U.hi}}, {{S1}}, {x1}, {y1}, {{S1}}", ["U.hi", "S1"], [x1, y1, "S1"])
    c.add("v_mad_u64_u32 {W}, {S2}, -1, 1, {U}", ["W", "S2"], ["U"])
    c.add("s_or_b64 {S1}, {S1}, {S2}", ["S1"], ["S1", "S2"])
    c.add(f"v_cndmask_b32 {{%X{tag}lo}}, {{U.lo}}, {{W.lo}}, {{S1}}", [f"%X{tag}lo"],
          ["U.lo", "W.lo", "S1"])
    c.add(f"v_cndmask_b32 {{%X{tag}hi}}, {{U.hi}}, {{W.hi}}, {{S1}}", [f"%X{tag}hi"],
          ["U.hi", "W.hi", "S1"])


def mac_chain(tag: str, base: int) -> Chain:
    """p' = p + montmul(x, k); x shared by both chains of a block."""
    c = Chain(tag, base)
    montmul(c, "%x0", "%x1", f"%k{tag}0", f"%k{tag}1")
    p0, p1 = f"%p{tag}0", f"%p{tag}1"
    c.add(f"v_add_co_u32 {{U.lo}}, {{S1}}, {p0}, {{A.lo}}", ["U.lo", "S1"], [p0, "A.lo"])
    c.add(f"v_addc_co_u32 {{U.hi}}, {{S1}}, {p1}, {{A.hi}}, {{S1}}", ["U.hi", "S1"], [p1, "A.hi", "S1"])
    c.add("v_cndmask_b32 {T1}, 0, -1, {S1}", ["T1"], ["S1"])
    c.add(f"v_mad_u64_u32 {{%P{tag}}}, {{D}}, {{T1}}, 1, {{U}}", [f"%P{tag}"], ["T1", "U"])
    return c


def r4bytes_chain(tag: str, base: int) -> Chain:
    """Stages 0 and 1 of the forward transform for one group of byte digits (a, b, c, d) at
    registers (g, g + 4, g + 8, g + 12): a, b, c, d are byte g of the shared dwords w0..w3,
    placed by v_perm_b32 (selector operands %<x><tag>: a, c16 = c << 16, b24 = b << 24,
    d8 = d << 8, b8 = b << 8, d24 = d << 24) into the chain's own temporaries. With the stage
    twiddles 2^48, -2^24 and -2^72 = 2^8 - 2^40 (mod P):
      A = a + 2^48 c = (c16 : a),  C = a - 2^48 c + P = (~c16 : a + 1),
      U = 2^24 b + 2^72 d = (d8 : b24) - (0 : d8),  V = 2^72 b + 2^24 d = (b8 : d24) - (0 : b8)
      (both in [0, 2^48)); outputs y0 = A - U (+P on borrow), y1 = A + U (< 2^57),
      y2 = C - V (C > 2^63 > V), y3 = C + V (+EPS on carry)."""
    c = Chain(tag, base)
    w0, w1, w2, w3 = "%w0", "%w1", "%w2", "%w3"
    sa, sc, sbh, sdl, sbl, sdh = (f"%s{n}{tag}" for n in ("a", "c", "bh", "dl", "bl", "dh"))
    c.add(f"v_perm_b32 {{W.lo}}, 0, {w1}, {{{sbh}}}", ["W.lo"], [w1, sbh])
    c.add(f"v_perm_b32 {{W.hi}}, 0, {w3}, {{{sdl}}}", ["W.hi"], [w3, sdl])
    c.add("v_sub_co_u32 {U.lo}, {S1}, {W.lo}, {W.hi}", ["U.lo", "S1"], ["W.lo", "W.hi"])
    c.add("v_subb_co_u32 {U.hi}, {S1}, {W.hi}, 0, {S1}", ["U.hi", "S1"], ["W.hi", "S1"])
    c.add(f"v_perm_b32 {{A.lo}}, 0, {w3}, {{{sdh}}}", ["A.lo"], [w3, sdh])
    c.add(f"v_perm_b32 {{A.hi}}, 0, {w1}, {{{sbl}}}", ["A.hi"], [w1, sbl])
    c.add("v_sub_co_u32 {V.lo}, {S2}, {A.lo}, {A.hi}", ["V.lo", "S2"], ["A.lo", "A.hi"])
    c.add("v_subb_co_u32 {V.hi}, {S2}, {A.hi}, 0, {S2}", ["V.hi", "S2"], ["A.hi", "S2"])
    c.add(f"v_perm_b32 {{T1}}, 0, {w0}, {{{sa}}}", ["T1"], [w0, sa])
    c.add(f"v_perm_b32 {{T2}}, 0, {w2}, {{{sc}}}", ["T2"], [w2, sc])
    c.add("v_add_u32 {Z.lo}, 1, {T1}", ["Z.lo"], ["T1"])
    c.add("v_not_b32 {Z.hi}, {T2}", ["Z.hi"], ["T2"])
    # y1 = A + U and y0 = A - U
    c.add(f"v_add_co_u32 {{%Y1{tag}lo}}, {{S1}}, {{T1}}, {{U.lo}}", [f"%Y1{tag}lo", "S1"], ["T1", "U.lo"])
    c.add(f"v_addc_co_u32 {{%Y1{tag}hi}}, {{S1}}, {{T2}}, {{U.hi}}, {{S1}}", [f"%Y1{tag}hi", "S1"],
          ["T2", "U.hi", "S1"])
    c.add("v_sub_co_u32 {W.lo}, {S2}, {T1}, {U.lo}", ["W.lo", "S2"], ["T1", "U.lo"])
    c.add("v_subb_co_u32 {W.hi}, {S2}, {T2}, {U.hi}, {S2}", ["W.hi", "S2"], ["T2", "U.hi", "S2"])
    c.add("v_cndmask_b32 {T1}, 0, -3, {S2}", ["T1"], ["S2"])
    c.add(f"v_mad_i64_i32 {{%Y0{tag}}}, {{D}}, {{T1}}, {{%k55}}, {{W}}", [f"%Y0{tag}"], ["T1", "%k55", "W"])
    # y2 = C - V and y3 = C + V
    c.add(f"v_sub_co_u32 {{%Y2{tag}lo}}, {{S1}}, {{Z.lo}}, {{V.lo}}", [f"%Y2{tag}lo", "S1"], ["Z.lo", "V.lo"])
    c.add(f"v_subb_co_u32 {{%Y2{tag}hi}}, {{S1}}, {{Z.hi}}, {{V.hi}}, {{S1}}", [f"%Y2{tag}hi", "S1"],
          ["Z.hi", "V.hi", "S1"])
    c.add("v_add_co_u32 {A.lo}, {S2}, {Z.lo}, {V.lo}", ["A.lo", "S2"], ["Z.lo", "V.lo"])
    c.add("v_addc_co_u32 {A.hi}, {S2}, {Z.hi}, {V.hi}, {S2}", ["A.hi", "S2"], ["Z.hi", "V.hi", "S2"])
    c.add("v_cndmask_b32 {T2}, 0, -1, {S2}", ["T2"], ["S2"])
    c.add(f"v_mad_u64_u32 {{%Y3{tag}}}, {{D}}, {{T2}}, 1, {{A}}", [f"%Y3{tag}"], ["T2", "A"])
    return c


def mul_chain(tag: str, base: int) -> Chain:
    """r = montmul(x, w), canonical; x in [0, 2^64), w < P (Montgomery constant form)."""
    c = Chain(tag, base)
    montmul(c, f"%x{tag}0", f"%x{tag}1", f"%w{tag}0", f"%w{tag}1", f"%R{tag}")
    return c


def schedule(chains: list[Chain], early_clobber: bool = False) -> list[str]:
    """List-schedule the chains; returns instruction lines with physical registers.
    early_clobber: the outputs are early-clobber operands (never share a register with an
    input), so they may be written before the last input read."""
    # nodes: (chain index, instr index)
    nodes = []
    for ci, c in enumerate(chains):
        for ii, (text, defs, uses) in enumerate(c.ins):
            nodes.append((ci, ii))
    # per-node dependence list: (pred node, min distance); soft: preferred distances
    # (result latency of v_mad_u64_u32, whose dependants stall when issued sooner)
    deps = {n: [] for n in nodes}
    soft = {n: [] for n in nodes}

    def atoms(ci, syms):
        out = []
        for s in syms:
            s = s.strip("{}")
            if s.startswith("%") or s == "D":
                out.append(s if s.startswith("%") else "D")
            else:
                out += [f"{ci}:{a}" for a in halves(s)]
        return out

    # With SHARED_Z every Z access of chain k comes after every Z access of chain k-1
    # (the chains' zero-extended products run one after the other).
    def uses_z(ins):
        return any(x.strip("{}").startswith("Z") for x in ins[1] + ins[2])

    for ci, c in enumerate(chains):
        for j, (tj, dj, uj) in enumerate(c.ins):
            aj_def, aj_use = atoms(ci, dj), atoms(ci, uj)
            for i in range(j):
                ti, di, ui = c.ins[i]
                ai_def, ai_use = atoms(ci, di), atoms(ci, ui)
                for r in set(ai_def) & set(aj_use):
                    if r == "D":
                        continue
                    sg = r.split(":")[-1].startswith("S")
                    valu_writer = not ti.startswith("s_")
                    valu_reader = not tj.startswith("s_")
                    dist = 3 if (sg and valu_writer and valu_reader) else 1
                    deps[(ci, j)].append(((ci, i), dist))
                    if not sg and ti.startswith("v_mad_u64_u32"):
                        soft[(ci, j)].append(((ci, i), MAD_LATENCY))
                for r in (set(ai_use) & set(aj_def)) | (set(ai_def) & set(aj_def)):
                    if r == "D":
                        continue
                    deps[(ci, j)].append(((ci, i), 1))
    for ci in range(1, len(chains) if SHARED_Z else 1):
        prev = [(ci - 1, i) for i, ins in enumerate(chains[ci - 1].ins) if uses_z(ins)]
        for j, ins in enumerate(chains[ci].ins):
            if uses_z(ins):
                deps[(ci, j)] += [(p, 1) for p in prev]
    # outputs (%X/%Y/%P defs) only after every read of any input operand
    in_readers = [] if early_clobber else [n for n in nodes
                  if any(u.startswith("%") and u != "%k55"
                         for u in chains[n[0]].ins[n[1]][2])]
    for n in nodes:
        d = chains[n[0]].ins[n[1]][1]
        if any(x.startswith("%") for x in d):
            for r in in_readers:
                if r != n:
                    deps[n].append((r, 1))
    # critical-path priority (with the soft latencies)
    succ = {n: [] for n in nodes}
    for n, ds in deps.items():
        for p, d in ds:
            succ[p].append((n, d))
        for p, d in soft[n]:
            succ[p].append((n, d))
    prio = {}

    def height(n):
        if n in prio:
            return prio[n]
        h = 1 + max((d + height(s) - 1 for s, d in succ[n]), default=0)
        prio[n] = h
        return h

    for n in nodes:
        height(n)
    slot = {}
    out = []
    t = 0
    remaining = set(nodes)
    while remaining:
        ready = [n for n in remaining
                 if all(p in slot and slot[p] + d <= t for p, d in deps[n])]
        if not ready:
            out.append(("s_nop 0", None))
            t += 1
            continue
        def soft_ready(n):
            return all(slot[p] + d <= t for p, d in soft[n])

        n = max(ready, key=lambda n: (soft_ready(n), prio[n], -n[1], -n[0]))
        slot[n] = t
        remaining.discard(n)
        out.append((chains[n[0]].ins[n[1]][0], n[0]))
        t += 1
    lines = []
    for text, ci in out:
        if ci is None:
            lines.append(text)
            continue
        c = chains[ci]

        def sub(m):
            s = m.group(1)
            if s.startswith("%"):
                return "{" + s + "}"
            if s == "D":
                return "{%dummy}"
            if s in ("S1", "S2"):
                return "{%" + f"s{s[1]}{c.tag}" + "}"
            return c.reg(s)

        lines.append(re.sub(r"\{([^}]+)\}", sub, text))
    return lines


def verify(lines: list[str], n_out: int, vgpr_inputs: set[int], early_clobber: bool = False) -> None:
    """Static checks of one emitted block (operands already numbered):
    * a VALU write of an SGPR operand and a later VALU read of it as carry-in or mask
      are at least 3 instructions apart (two wait states on gfx950);
    * no output operand (%0..%n_out-1) is written before the last read of a VGPR input
      operand, so outputs may share registers with inputs;
    * every fixed temporary is written in the block before it is read."""
    def ops(ln):
        mnem, _, rest = ln.partition(" ")
        args = [a.strip() for a in rest.split(",")] if rest else []
        return mnem, args

    sgpr_write = {}
    last_in_read = -1
    first_out_write = None
    written = set()
    for t, ln in enumerate(lines):
        mnem, args = ops(ln)
        if mnem == "s_nop":
            continue
        dsts, srcs = [], args
        if mnem.startswith("v_") or mnem.startswith("s_"):
            if mnem in ("v_add_co_u32", "v_sub_co_u32", "v_addc_co_u32", "v_subb_co_u32",
                        "v_mad_u64_u32", "v_mad_i64_i32"):
                dsts, srcs = args[:2], args[2:]
            else:
                dsts, srcs = args[:1], args[1:]
        for a in srcs:
            m = re.fullmatch(r"%(\d+)", a)
            if m:
                k = int(m.group(1))
                if k in vgpr_inputs:
                    last_in_read = t
                if a in sgpr_write and mnem.startswith("v_"):
                    assert t - sgpr_write[a] >= 3, f"SGPR hazard at {t}: {ln}"
            vm = re.fullmatch(r"v\[?(\d+)(?::(\d+)\])?", a)
            if vm:
                lo = int(vm.group(1))
                hi = int(vm.group(2) or lo)
                for r in range(lo, hi + 1):
                    assert r in written, f"read of unwritten v{r} at {t}: {ln}"
        for a in dsts:
            m = re.fullmatch(r"%(\d+)", a)
            if m:
                k = int(m.group(1))
                if k < n_out and first_out_write is None:
                    first_out_write = t
                if mnem.startswith("v_"):
                    sgpr_write[a] = t
                else:
                    sgpr_write.pop(a, None)
            vm = re.fullmatch(r"v\[?(\d+)(?::(\d+)\])?", a)
            if vm:
                lo = int(vm.group(1))
                hi = int(vm.group(2) or lo)
                written.update(range(lo, hi + 1))
    assert early_clobber or first_out_write is None or first_out_write > last_in_read, \
        "output before input read"


def emit(name: str, chains: list[Chain], outputs: list[tuple[str, str, str]],
         inputs: list[tuple[str, str, str]], doc: str, early_clobber: bool = False) -> str:
    """outputs/inputs: (symbol, constraint, C expression). early_clobber: the outputs'
    constraints are "=&v" and may be written before the last input read."""
    # the zero half of the zero-extension pair Z, for the chains that read Z at all
    def reads_z(c):
        """Whether the chain reads Z's zero half before writing all of Z."""
        for _, defs, uses in c.ins:
            if any(u.strip("{}") in ("Z", "Z.hi") for u in uses):
                return True
            if any(d.strip("{}") in ("Z", "Z.hi") for d in defs):
                return False
        return False
    zeros = ([TEMP_BASE + 9] if any(reads_z(c) for c in chains) else []) if SHARED_Z else \
        [c.base + 9 for c in chains if reads_z(c)]
    lines = [f"v_mov_b32 v{z}, 0" for z in zeros] + schedule(chains, early_clobber)
    sg = []
    for c in chains:
        sg += [(f"s1{c.tag}", "=&s", f"s1{c.tag}"), (f"s2{c.tag}", "=&s", f"s2{c.tag}")]
    ops = outputs + [("dummy", "=&s", "dummy")] + sg + inputs
    index = {sym: i for i, (sym, _, _) in enumerate(ops)}
    body = []
    for ln in lines:
        ln = re.sub(r"\{%([A-Za-z0-9]+)\}", lambda m: f"%{index[m.group(1)]}", ln)
        ln = re.sub(r"%([A-Za-z][A-Za-z0-9]*)", lambda m: f"%{index[m.group(1)]}", ln)
        body.append(f'      "{ln}\\n"')
    outs = ", ".join(f'"{c}"({e})' for _, c, e in outputs + [("dummy", "=&s", "dummy")] + sg)
    ins = ", ".join(f'"{c}"({e})' for _, c, e in inputs)
    clob = ", ".join(f'"v{r}"' for r in range(TEMP_BASE, 256))
    numbered = [re.sub(r"%([A-Za-z][A-Za-z0-9]*)", lambda m: f"%{index[m.group(1)]}",
                       re.sub(r"\{%([A-Za-z0-9]+)\}", lambda m: f"%{index[m.group(1)]}", ln))
                for ln in lines]
    verify(numbered, len(outputs), {index[sym] for sym, c, _ in inputs if c == "v"}, early_clobber)
    n_nop = sum(1 for ln in lines if ln.startswith("s_nop"))
    return (f"// {doc}\n// {len(lines)} instructions, {n_nop} of them s_nop, for "
            f"{len(chains)} chains.\n"
            f"#define HB_GL_{name.upper()}_ASM \\\n"
            "  asm( \\\n" + " \\\n".join(body) + " \\\n"
            f"      : {outs} \\\n      : {ins} \\\n      : {clob}, \"scc\")\n")


def main():
    parts = ["// Generated by gen_gl_asm.py — do not edit.\n"
             "// Each macro expects the variables its operand lists name to be in scope.\n"]
    for wkind, wc in (("v", "v"), ("s", "s")):
        a, b = ct_chain("a", TEMP_BASE, wkind), ct_chain("b", TEMP_BASE + 12, wkind)
        outputs = [("Xa", "=v", "Xa"), ("Ya", "=v", "Ya"), ("Xb", "=v", "Xb"), ("Yb", "=v", "Yb")]
        inputs = []
        for t in "ab":
            inputs += [(f"x{t}0", "v", f"(u32)x{t}"), (f"x{t}1", "v", f"(u32)(x{t} >> 32)"),
                       (f"y{t}0", "v", f"(u32)y{t}"), (f"y{t}1", "v", f"(u32)(y{t} >> 32)"),
                       (f"{wkind}w{t}0", wc, f"(u32)w{t}"),
                       (f"{wkind}w{t}1", wc, f"(u32)(w{t} >> 32)")]
        inputs.append(("k55", "s", "0x55555555u"))
        parts.append(emit(f"ct2_{wkind}", [a, b], outputs, inputs,
                          f"Two CT butterflies (x, y) <- (x + w*y, x - w*y), twiddles in "
                          f"{'SGPRs' if wkind == 's' else 'VGPRs'}."))
        a, b = gs_chain("a", TEMP_BASE, wkind), gs_chain("b", TEMP_BASE + 12, wkind)
        outputs = [("Xalo", "=v", "Xalo"), ("Xahi", "=v", "Xahi"), ("Ya", "=v", "Ya"),
                   ("Xblo", "=v", "Xblo"), ("Xbhi", "=v", "Xbhi"), ("Yb", "=v", "Yb")]
        parts.append(emit(f"gs2_{wkind}", [a, b], outputs, inputs,
                          f"Two GS butterflies (x, y) <- (x + y, (x - y)*w) on canonical "
                          f"inputs, twiddles in {'SGPRs' if wkind == 's' else 'VGPRs'}."))
    # power-of-two twiddles: one macro per pair of multiplier types
    for ta in ("A", "BL", "BH"):
        for tb in ("A", "BL", "BH"):
            inputs = []
            for t in "ab":
                inputs += [(f"x{t}0", "v", f"(u32)x{t}"), (f"x{t}1", "v", f"(u32)(x{t} >> 32)"),
                           (f"y{t}0", "v", f"(u32)y{t}"), (f"y{t}1", "v", f"(u32)(y{t} >> 32)"),
                           (f"s{t}1", "n", f"S{t.upper()}1"), (f"s{t}2", "n", f"S{t.upper()}2")]
            inputs.append(("k55", "s", "0x55555555u"))
            a, b = ct_p2_chain("a", TEMP_BASE, ta), ct_p2_chain("b", TEMP_BASE + 12, tb)
            outputs = [("Xa", "=v", "Xa"), ("Ya", "=v", "Ya"), ("Xb", "=v", "Xb"),
                       ("Yb", "=v", "Yb")]
            parts.append(emit(f"ctp2_{ta}_{tb}", [a, b], outputs, inputs,
                              f"Two CT butterflies with power-of-two twiddles ({ta}, {tb}): "
                              "(x, y) <- (x + t, x - t), t = y * 2^e by shifts."))
            a, b = gs_p2_chain("a", TEMP_BASE, ta), gs_p2_chain("b", TEMP_BASE + 12, tb)
            outputs = [("Xalo", "=v", "Xalo"), ("Xahi", "=v", "Xahi"), ("Ya", "=v", "Ya"),
                       ("Xblo", "=v", "Xblo"), ("Xbhi", "=v", "Xbhi"), ("Yb", "=v", "Yb")]
            parts.append(emit(f"gsp2_{ta}_{tb}", [a, b], outputs, inputs,
                              f"Two GS butterflies with power-of-two twiddles ({ta}, {tb}) on "
                              "canonical inputs: (x, y) <- (x + y, (x - y) * 2^e)."))
    # CT pairs with a twiddle of 1 in at least one chain (the twisted forward transform)
    for ta in ("A", "BL", "BH", "TN", "TC"):
        for tb in ("A", "BL", "BH", "TN", "TC"):
            if ta not in ("TN", "TC") and tb not in ("TN", "TC"):
                continue
            inputs = []
            for t, typ in (("a", ta), ("b", tb)):
                inputs += [(f"x{t}0", "v", f"(u32)x{t}"), (f"x{t}1", "v", f"(u32)(x{t} >> 32)"),
                           (f"y{t}0", "v", f"(u32)y{t}"), (f"y{t}1", "v", f"(u32)(y{t} >> 32)")]
                if typ not in ("TN", "TC"):
                    inputs += [(f"s{t}1", "n", f"S{t.upper()}1"), (f"s{t}2", "n", f"S{t.upper()}2")]
            inputs.append(("k55", "s", "0x55555555u"))
            a, b = ct_any_chain("a", TEMP_BASE, ta), ct_any_chain("b", TEMP_BASE + 12, tb)
            outputs = [("Xa", "=v", "Xa"), ("Ya", "=v", "Ya"), ("Xb", "=v", "Xb"),
                       ("Yb", "=v", "Yb")]
            parts.append(emit(f"ctp2_{ta}_{tb}", [a, b], outputs, inputs,
                              f"Two CT butterflies ({ta}, {tb}): (x, y) <- (x + t, x - t), t = y "
                              "for TC (y canonical) and TN (y any word), else y * 2^e by shifts."))
    a, b = mac_chain("a", TEMP_BASE), mac_chain("b", TEMP_BASE + 12)
    outputs = [("Pa", "=v", "Pa"), ("Pb", "=v", "Pb")]
    inputs = [("x0", "v", "(u32)x"), ("x1", "v", "(u32)(x >> 32)")]
    for t in "ab":
        inputs += [(f"k{t}0", "v", f"(u32)k{t}"), (f"k{t}1", "v", f"(u32)(k{t} >> 32)"),
                   (f"p{t}0", "v", f"(u32)p{t}"), (f"p{t}1", "v", f"(u32)(p{t} >> 32)")]
    inputs.append(("k55", "s", "0x55555555u"))
    parts.append(emit("mac2", [a, b], outputs, inputs,
                      "(pa, pb) += (x * ka, x * kb), keys in Montgomery constant form."))
    a, b = r4bytes_chain("a", TEMP_BASE), r4bytes_chain("b", TEMP_BASE + 12)
    outputs = []
    for t in "ab":
        outputs += [(f"Y0{t}", "=&v", f"Y0{t}"), (f"Y1{t}lo", "=&v", f"Y1{t}lo"),
                    (f"Y1{t}hi", "=&v", f"Y1{t}hi"), (f"Y2{t}lo", "=&v", f"Y2{t}lo"),
                    (f"Y2{t}hi", "=&v", f"Y2{t}hi"), (f"Y3{t}", "=&v", f"Y3{t}")]
    inputs = [(f"w{m}", "v", f"w[{m}]") for m in range(4)]
    for t, g in (("a", "G"), ("b", "(G + 1)")):
        inputs += [(f"sa{t}", "s", f"0x0C0C0C00u | {g}"), (f"sc{t}", "s", f"0x0C000C0Cu | ({g} << 16)"),
                   (f"sbh{t}", "s", f"0x000C0C0Cu | ({g} << 24)"),
                   (f"sdl{t}", "s", f"0x0C0C000Cu | ({g} << 8)"),
                   (f"sbl{t}", "s", f"0x0C0C000Cu | ({g} << 8)"),
                   (f"sdh{t}", "s", f"0x000C0C0Cu | ({g} << 24)")]
    inputs.append(("k55", "s", "0x55555555u"))
    parts.append(emit("r4bytes2", [a, b], outputs, inputs,
                      "Forward stages 0 and 1 of byte-digit groups G and G + 1 (r4bytes_chain): "
                      "y0..y3 in [0, 2^64) from the digit dwords w[0..3].", early_clobber=True))
    a, b = mul_chain("a", TEMP_BASE), mul_chain("b", TEMP_BASE + 12)
    outputs = [("Ra", "=v", "Ra"), ("Rb", "=v", "Rb")]
    inputs = []
    for t in "ab":
        inputs += [(f"x{t}0", "v", f"(u32)x{t}"), (f"x{t}1", "v", f"(u32)(x{t} >> 32)"),
                   (f"w{t}0", "v", f"(u32)w{t}"), (f"w{t}1", "v", f"(u32)(w{t} >> 32)")]
    inputs.append(("k55", "s", "0x55555555u"))
    parts.append(emit("mul2", [a, b], outputs, inputs,
                      "(ra, rb) <- (xa * wa, xb * wb), constants in Montgomery form; canonical."))
    text = "\n".join(parts)
    if len(sys.argv) > 1 and sys.argv[1] == "--check":
        return 0 if open(OUT).read() == text else 1
    with open(OUT, "w") as f:
        f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())

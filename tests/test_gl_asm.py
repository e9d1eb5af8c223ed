"""CPU checks of the hand-scheduled Goldilocks sequences (hasteboots_amd/csrc/hb_gl_asm.inc).

The generated file must match its generator, and every macro is run through a one-lane
interpreter of the eleven gfx950 instructions it uses, against Python big-integer
arithmetic: CT and GS butterflies and the MAC with Montgomery-form constants
(c stands for c * 2^-64 mod P), on random and edge words.
"""
import random
import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "hasteboots_amd" / "csrc"
INC = CSRC / "hb_gl_asm.inc"
P = 2**64 - 2**32 + 1
EPS = 2**32 - 1
M32, M64 = 2**32 - 1, 2**64 - 1
RINV = pow(2**64, P - 2, P)
EDGE = [0, 1, 2, EPS, 2**32, 2**63, P - 2, P - 1, P, P + 1, M64 - 1, M64, 2**64 - 2**32]


def parse_macros():
    text = INC.read_text()
    macros = {}
    for m in re.finditer(r"#define (HB_GL_\w+_ASM) \\\n  asm\( \\\n(.*?)\)\n", text, re.S):
        body = m.group(2)
        lines = re.findall(r'^\s+"(.*?)\\n" \\$', body, re.M)
        outs, ins = re.findall(r"^\s+: (.*?) \\$", body, re.M)[:2]
        ops = re.findall(r'"([=&a-z]+)"\(([^"]*?)\)(?:, |$)', outs + ", " + ins + ", ")
        macros[m.group(1)] = (lines, ops)
    return macros


def c_eval(expr, env):
    if expr in env:
        return env[expr]
    e = re.sub(r"(0x[0-9A-Fa-f]+)u\b", r"\1", expr)
    e = re.sub(r"\(u32\)\(([^()]*)\)", r"u32(\1)", e)
    e = re.sub(r"\(u32\)(\w+)", r"u32(\1)", e)
    return eval(e, {"u32": lambda v: v & M32}, env)


class Lane:
    """One lane of the VALU: 32-bit VGPRs, SGPR pairs reduced to this lane's bit."""

    def __init__(self, operands):
        self.v = {}
        self.op = dict(operands)  # operand index -> value (64-bit for pairs)

    def rd32(self, a):
        if a.startswith("%"):
            return self.op[int(a[1:])] & M32
        if re.fullmatch(r"v\d+", a):
            return self.v[int(a[1:])]
        return int(a) & M32

    def rd64(self, a):
        if a.startswith("%"):
            return self.op[int(a[1:])] & M64
        m = re.fullmatch(r"v\[(\d+):(\d+)\]", a)
        if m:
            lo = int(m.group(1))
            return self.v[lo] | (self.v[lo + 1] << 32)
        return int(a) & M64

    def wr32(self, a, x):
        if a.startswith("%"):
            self.op[int(a[1:])] = x & M32
        else:
            self.v[int(a[1:])] = x & M32

    def wr64(self, a, x):
        if a.startswith("%"):
            self.op[int(a[1:])] = x & M64
        else:
            lo = int(re.fullmatch(r"v\[(\d+):\d+\]", a).group(1))
            self.v[lo], self.v[lo + 1] = x & M32, (x >> 32) & M32

    def bit(self, a):
        return self.op[int(a[1:])] & 1

    def run(self, lines):
        for ln in lines:
            mn, _, rest = ln.partition(" ")
            a = [x.strip() for x in rest.split(",")] if rest else []
            if mn == "s_nop":
                continue
            if mn == "v_mov_b32":
                self.wr32(a[0], self.rd32(a[1]))
            elif mn == "v_mov_b64":
                self.wr64(a[0], self.rd64(a[1]))
            elif mn == "v_mad_u64_u32":
                r = self.rd32(a[2]) * self.rd32(a[3]) + self.rd64(a[4])
                self.wr64(a[0], r)
                self.op[int(a[1][1:])] = r >> 64
            elif mn == "v_mad_i64_i32":
                def s32(x):
                    return x - 2**32 if x >> 31 else x
                c = self.rd64(a[4])
                c = c - 2**64 if c >> 63 else c
                self.wr64(a[0], s32(self.rd32(a[2])) * s32(self.rd32(a[3])) + c)
            elif mn in ("v_add_co_u32", "v_addc_co_u32"):
                assert len(a) == (5 if mn == "v_addc_co_u32" else 4), ln
                r = self.rd32(a[2]) + self.rd32(a[3]) + (self.bit(a[4]) if len(a) > 4 else 0)
                self.wr32(a[0], r)
                self.op[int(a[1][1:])] = r >> 32
            elif mn in ("v_sub_co_u32", "v_subb_co_u32"):
                assert len(a) == (5 if mn == "v_subb_co_u32" else 4), ln
                r = self.rd32(a[2]) - self.rd32(a[3]) - (self.bit(a[4]) if len(a) > 4 else 0)
                self.wr32(a[0], r)
                self.op[int(a[1][1:])] = 1 if r < 0 else 0
            elif mn == "v_cndmask_b32":
                self.wr32(a[0], self.rd32(a[2]) if self.bit(a[3]) else self.rd32(a[1]))
            elif mn == "v_lshlrev_b32":
                self.wr32(a[0], self.rd32(a[2]) << (self.rd32(a[1]) & 31))
            elif mn == "v_lshrrev_b32":
                self.wr32(a[0], self.rd32(a[2]) >> (self.rd32(a[1]) & 31))
            elif mn == "v_lshrrev_b64":
                self.wr64(a[0], self.rd64(a[2]) >> (self.rd32(a[1]) & 63))
            elif mn == "v_alignbit_b32":
                v = (self.rd32(a[1]) << 32) | self.rd32(a[2])
                self.wr32(a[0], v >> (self.rd32(a[3]) & 31))
            elif mn == "v_perm_b32":
                s0, s1, sel = self.rd32(a[1]), self.rd32(a[2]), self.rd32(a[3])
                pool = [(s1 >> (8 * i)) & 0xFF for i in range(4)] + [(s0 >> (8 * i)) & 0xFF for i in range(4)]
                r = 0
                for i in range(4):
                    b = (sel >> (8 * i)) & 0xFF
                    assert b < 8 or b == 0x0C, hex(sel)
                    r |= (pool[b] if b < 8 else 0) << (8 * i)
                self.wr32(a[0], r)
            elif mn == "v_not_b32":
                self.wr32(a[0], ~self.rd32(a[1]))
            elif mn == "v_add_u32":
                self.wr32(a[0], self.rd32(a[1]) + self.rd32(a[2]))
            elif mn == "s_or_b64":
                self.op[int(a[0][1:])] = self.bit(a[1]) | self.bit(a[2])
            else:
                raise AssertionError(f"unknown instruction {ln}")


def run_macro(macro, env):
    lines, ops = macro
    vals = {}
    for i, (con, expr) in enumerate(ops):
        if "=" not in con:
            vals[i] = c_eval(expr, env)
    lane = Lane(vals)
    lane.run(lines)
    return {expr: lane.op[i] for i, (con, expr) in enumerate(ops) if con in ("=v", "=&v")}


def mont(y, c):
    return y * c * RINV % P


def samples(rng, n, canonical=False):
    top = P if canonical else 2**64
    out = [rng.randrange(top) for _ in range(n)]
    out += [e for e in EDGE if e < top]
    return out


def test_generated_file_is_current():
    r = subprocess.run([sys.executable, str(CSRC / "gen_gl_asm.py"), "--check"])
    assert r.returncode == 0, "hb_gl_asm.inc differs from gen_gl_asm.py output"


@pytest.mark.parametrize("name", ["HB_GL_CT2_V_ASM", "HB_GL_CT2_S_ASM"])
def test_ct_pair(name):
    mac = parse_macros()[name]
    rng = random.Random(1)
    xs, ys, ws = samples(rng, 300), samples(rng, 300), samples(rng, 300, canonical=True)
    for i in range(len(xs)):
        env = dict(xa=xs[i], ya=ys[i], wa=ws[i % len(ws)], xb=ys[-1 - i], yb=xs[i],
                   wb=ws[(7 * i) % len(ws)])
        out = run_macro(mac, env)
        for t in "ab":
            x, y, w = env["x" + t], env["y" + t], env["w" + t]
            tt = mont(y, w)
            assert out["X" + t] % P == (x + tt) % P
            assert out["Y" + t] % P == (x - tt) % P


@pytest.mark.parametrize("name", ["HB_GL_GS2_V_ASM", "HB_GL_GS2_S_ASM"])
def test_gs_pair(name):
    mac = parse_macros()[name]
    rng = random.Random(2)
    xs, ys, ws = (samples(rng, 300, canonical=True) for _ in range(3))
    for i in range(len(xs)):
        env = dict(xa=xs[i], ya=ys[i], wa=ws[i], xb=ys[-1 - i], yb=xs[-1 - i],
                   wb=ws[(5 * i) % len(ws)])
        out = run_macro(mac, env)
        for t in "ab":
            x, y, w = env["x" + t], env["y" + t], env["w" + t]
            sx = out[f"X{t}lo"] | (out[f"X{t}hi"] << 32)
            assert sx == (x + y) % P  # canonical sum
            assert out["Y" + t] == mont((x - y) % P, w)  # canonical product


def test_mac_pair():
    mac = parse_macros()["HB_GL_MAC2_ASM"]
    rng = random.Random(3)
    xs, ps = samples(rng, 300), samples(rng, 300)
    ks = samples(rng, 300, canonical=True)
    for i in range(len(xs)):
        env = dict(x=xs[i], ka=ks[i % len(ks)], kb=ks[-1 - i % len(ks)], pa=ps[i], pb=ps[-1 - i])
        out = run_macro(mac, env)
        assert out["Pa"] % P == (env["pa"] + mont(env["x"], env["ka"])) % P
        assert out["Pb"] % P == (env["pb"] + mont(env["x"], env["kb"])) % P


def test_mul_pair():
    mul = parse_macros()["HB_GL_MUL2_ASM"]
    rng = random.Random(4)
    xs = samples(rng, 300)
    ws = samples(rng, 300, canonical=True)
    for i in range(len(xs)):
        env = dict(xa=xs[i], wa=ws[i % len(ws)], xb=xs[-1 - i], wb=ws[(7 * i) % len(ws)])
        out = run_macro(mul, env)
        assert out["Ra"] == mont(env["xa"], env["wa"])  # canonical product
        assert out["Rb"] == mont(env["xb"], env["wb"])


def test_byte_digit_stages():
    """HB_GL_R4BYTES2_ASM: forward stages 0 and 1 of a group of byte digits (a, b, c, d) at
    registers (g, g+4, g+8, g+12), against the transform's own twiddles psi^512, psi^256,
    psi^768 (psi of order 2048), on random and extreme bytes."""
    mac = parse_macros()["HB_GL_R4BYTES2_ASM"]
    psi = pow(7, (P - 1) // 2048, P)
    w0, w1, w2 = pow(psi, 512, P), pow(psi, 256, P), pow(psi, 768, P)
    rng = random.Random(5)
    groups = [tuple(rng.randrange(256) for _ in range(4)) for _ in range(400)]
    groups += [(a, b, c, d) for a in (0, 255) for b in (0, 255) for c in (0, 255) for d in (0, 255)]
    for i in range(0, len(groups) - 1, 2):
        G = (i // 2) % 3  # groups G and G + 1: bytes G, G + 1 of the dwords
        w = [rng.randrange(2**32) for _ in range(4)]  # the other bytes are noise
        for t, (a, b, c, d) in zip("ab", groups[i:i + 2]):
            g = G + (t == "b")
            for m, v in enumerate((a, b, c, d)):
                w[m] = (w[m] & ~(0xFF << (8 * g))) | (v << (8 * g))
        out = run_macro(mac, dict(w=w, G=G))
        for t, (a, b, c, d) in zip("ab", groups[i:i + 2]):
            A, C = (a + w0 * c) % P, (a - w0 * c) % P
            B, D = (b + w0 * d) % P, (b - w0 * d) % P
            want = [(A + w1 * B) % P, (A - w1 * B) % P, (C + w2 * D) % P, (C - w2 * D) % P]
            got = [out[f"Y0{t}"], out[f"Y1{t}lo"] | (out[f"Y1{t}hi"] << 32),
                   out[f"Y2{t}lo"] | (out[f"Y2{t}hi"] << 32), out[f"Y3{t}"]]
            assert [g % P for g in got] == want, (a, b, c, d)


def _pow2_cases(typ):
    """(s1, s2, exponent e of the multiplier 2^e mod P) for each shift of a type."""
    if typ == "A":
        return [(k, 32 - k, k) for k in range(1, 32)]
    if typ == "BL":
        return [(j, 32 - j, j - 64) for j in range(1, 32)]
    return [(j - 32, 64 - j, j - 64) for j in range(33, 64)]


@pytest.mark.parametrize("ta", ["A", "BL", "BH", "TN", "TC"])
@pytest.mark.parametrize("tb", ["A", "BL", "BH", "TN", "TC"])
def test_ct_pairs_with_unit_twiddle(ta, tb):
    """CT pairs where a chain's twiddle is 1: TN takes any y word, TC a canonical y."""
    if "T" not in ta[0] + tb[0]:
        pytest.skip("no unit twiddle (test_pow2_pairs)")
    mac = parse_macros()[f"HB_GL_CTP2_{ta}_{tb}_ASM"]
    rng = random.Random(hash((ta, tb)) & 0xFFFF)
    cases = {t: ([(0, 0, 0)] if t in ("TN", "TC") else _pow2_cases(t)) for t in (ta, tb)}
    anyw, canw = samples(rng, 60), samples(rng, 60, canonical=True)
    for i in range(max(len(cases[ta]), len(cases[tb])) * 3 + 60):
        sa, sb = cases[ta][i % len(cases[ta])], cases[tb][(7 * i) % len(cases[tb])]
        ya = (canw if ta == "TC" else anyw)[(3 * i + 1) % 60]
        yb = (canw if tb == "TC" else anyw)[(11 * i + 3) % 60]
        env = dict(xa=anyw[i % 60], ya=ya, xb=anyw[(5 * i + 2) % 60], yb=yb,
                   SA1=sa[0], SA2=sa[1], SB1=sb[0], SB2=sb[1])
        out = run_macro(mac, env)
        for t, (_, _, e) in (("a", sa), ("b", sb)):
            x, y = env["x" + t], env["y" + t]
            m = pow(2, e % 192, P)
            assert out["X" + t] % P == (x + y * m) % P, (t, e)
            assert out["Y" + t] % P == (x - y * m) % P, (t, e)


def test_unit_twiddle_mutation_is_caught():
    """A TN chain that skipped the canonicalisation of y gives a wrong word for some y >= P
    (the single EPS correction of x + y cannot absorb two wraps): the edge words catch it."""
    lines, ops = parse_macros()["HB_GL_CTP2_TN_TN_ASM"]
    bad = [ln for ln in lines if not ln.startswith("v_cndmask_b32") or ", -1," in ln or ", 0, -" in ln]
    broken = (bad, ops)
    found = False
    for x in EDGE:
        for y in EDGE:
            env = dict(xa=x, ya=y, xb=x, yb=y, SA1=0, SA2=0, SB1=0, SB2=0)
            try:
                out = run_macro(broken, env)
            except KeyError:
                found = True
                break
            if out["Xa"] % P != (x + y) % P or out["Ya"] % P != (x - y) % P:
                found = True
                break
        if found:
            break
    assert found


@pytest.mark.parametrize("kind", ["CT", "GS"])
@pytest.mark.parametrize("ta", ["A", "BL", "BH"])
@pytest.mark.parametrize("tb", ["A", "BL", "BH"])
def test_pow2_pairs(kind, ta, tb):
    mac = parse_macros()[f"HB_GL_{kind}P2_{ta}_{tb}_ASM"]
    rng = random.Random(hash((kind, ta, tb)) & 0xFFFF)
    ca, cb = _pow2_cases(ta), _pow2_cases(tb)
    canon = kind == "GS"
    vals = samples(rng, 60, canonical=canon)
    for i in range(max(len(ca), len(cb)) * 3):
        sa, sb = ca[i % len(ca)], cb[(7 * i) % len(cb)]
        env = dict(xa=vals[i % len(vals)], ya=vals[(3 * i + 1) % len(vals)],
                   xb=vals[(5 * i + 2) % len(vals)], yb=vals[(11 * i + 3) % len(vals)],
                   SA1=sa[0], SA2=sa[1], SB1=sb[0], SB2=sb[1])
        out = run_macro(mac, env)
        for t, (_, _, e) in (("a", sa), ("b", sb)):
            x, y = env["x" + t], env["y" + t]
            m = pow(2, e % 192, P)
            if kind == "CT":
                assert out["X" + t] % P == (x + y * m) % P, (t, e)
                assert out["Y" + t] % P == (x - y * m) % P, (t, e)
            else:
                assert (out[f"X{t}lo"] | (out[f"X{t}hi"] << 32)) == (x + y) % P
                assert out["Y" + t] == (x - y) * m % P, (t, e)

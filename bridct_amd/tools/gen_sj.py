#!/usr/bin/env python3
"""Shao-Johnson butterfly generator for the MI355X BRiDCT kernels.

The reference ships its butterfly graphs only as frozen generated C (the generator it
names, tools/extract.py, is absent: src/generated/large.c:1, rect_generated.inc).  This
tool derives the graphs again from the published mathematics and emits CDNA4 HIP device
functions:

* Johnson & Frigo, "A modified split-radix FFT with fewer arithmetic operations",
  IEEE TSP 55(1), 2007: four mutually recursive transforms (plain, S, S2, S4) whose
  outputs are divided by the scale s_{N,k} (plan.c:39-47 restates the scale).
* Shao & Johnson, "Type-II/III DCT/DST algorithms with reduced number of arithmetic
  operations", Signal Processing 88(6), 2008: the DCT-II of length n as the Makhoul
  reordering (rect_generated.inc:87 idx8_0), a real S-scaled split-radix FFT and a
  tangent post-twiddle (post8_body, rect_generated.inc:89-96).  The resulting 1D core is
  F_n[k,j] = cos(pi k (2j+1) / (2n)) / s_{4n,k}; the inverse core is F_n^T
  (reference plan.c:38 comment, :96).

Graphs are built symbolically as linear DAGs with hash-consed nodes of the form
``a + r*b`` (an add, a subtract or one FMA).  Constant factors are deferred along the
graph and folded into the next FMA, which is how the JF scaled twiddles (1 - i tan) come
out; a factor still pending at an output becomes one multiply.  Inverse cores are the
exact transposes of the forward graphs (transposition principle), as the reference
inverts its graphs (col_t*, rect_generated.inc:110-114).

Every emitted graph is checked here in FP64 against its closed form before anything is
written; ``--check`` runs only the checks.

Outputs (all under bridct_amd/csrc/generated/):
  sj_codelets.h   HIP device functions: sjf_n / sjt_n (n = 8..64, real, in place) and
                  dft_N_m / dft_N_p (N = 2..32, complex split radix, sign -/+).
  sj_factors.h    host FP64 table d_n: sjf_n = diag(1/d_n) F_n (folded into normalization).
  sj_counts.json  operation counts per graph.
"""
import argparse
import json
import math
import os
import sys

import numpy as np

SNAP = 1e-12


def sj_scale(N, k):
    """Johnson-Frigo scale s_{N,k} (reference src/plan.c:39-47)."""
    if N <= 4:
        return 1.0
    q = N // 4
    k %= q
    a = 2.0 * math.pi * k / N
    return sj_scale(q, k) * (math.cos(a) if k <= N // 8 else math.sin(a))


def snap(c):
    if abs(c - 1.0) < SNAP:
        return 1.0
    if abs(c + 1.0) < SNAP:
        return -1.0
    return c


class Graph:
    """Linear DAG.  ops[i] is ('in', k), ('lin', a, b, r) meaning v[a] + r * v[b], or
    ('mul', a, f) meaning f * v[a].

    A handle is None (exact zero) or (c, i) meaning c * v[i]; c is a deferred factor.
    """

    def __init__(self, n_inputs):
        self.ops = []
        self.memo = {}
        self.inputs = [self._new(('in', i)) for i in range(n_inputs)]

    def _new(self, op):
        self.ops.append(op)
        return len(self.ops) - 1

    def inp(self, i):
        return (1.0, self.inputs[i])

    def _pair(self, t1, t2):
        (c1, a), (c2, b) = t1, t2
        if a > b:
            (c1, a), (c2, b) = (c2, b), (c1, a)
        # factor out the coefficient that leaves no pending multiply where possible
        if abs(abs(c1) - abs(c2)) <= SNAP * abs(c1) or snap(c1) in (1.0, -1.0):
            f, first, second, cs = c1, a, b, c2
        elif snap(c2) in (1.0, -1.0):
            f, first, second, cs = c2, b, a, c1
        else:
            f, first, second, cs = c1, a, b, c2
        r = snap(cs / f)
        key = (first, second, float('%.12e' % r))
        nid = self.memo.get(key)
        if nid is None:
            nid = self._new(('lin', first, second, r))
            self.memo[key] = nid
        return (snap(f), nid)

    def comb(self, terms):
        acc, big, order = {}, {}, []
        for s, h in terms:
            if h is None or s == 0.0:
                continue
            c, nid = h
            if nid not in acc:
                acc[nid], big[nid] = 0.0, 0.0
                order.append(nid)
            acc[nid] += s * c
            big[nid] = max(big[nid], abs(s * c))
        items = [(acc[n], n) for n in order if abs(acc[n]) > 1e-12 * big[n]]
        if not items:
            return None
        while len(items) > 2:
            items = [self._pair(items[0], items[1])] + items[2:]
        if len(items) == 1:
            return (snap(items[0][0]), items[0][1])
        return self._pair(items[0], items[1])

    def mul(self, i, f):
        """Node f * v[i]."""
        key = ('mul', i, float('%.12e' % f))
        nid = self.memo.get(key)
        if nid is None:
            nid = self._new(('mul', i, f))
            self.memo[key] = nid
        return nid

    def lin(self, a, b, r):
        """Node v[a] + r * v[b] (no factoring)."""
        key = (a, b, float('%.12e' % r))
        nid = self.memo.get(key)
        if nid is None:
            nid = self._new(('lin', a, b, r))
            self.memo[key] = nid
        return nid

    def live(self, outputs):
        seen = set()
        stack = [h[1] for h in outputs if h is not None]
        while stack:
            i = stack.pop()
            if i in seen:
                continue
            seen.add(i)
            op = self.ops[i]
            if op[0] == 'lin':
                stack.extend((op[1], op[2]))
            elif op[0] == 'mul':
                stack.append(op[1])
        return seen

    def counts(self, outputs):
        live = self.live(outputs)
        adds = fmas = inner_muls = 0
        for i in live:
            op = self.ops[i]
            if op[0] == 'lin':
                if op[3] in (1.0, -1.0):
                    adds += 1
                else:
                    fmas += 1
            elif op[0] == 'mul':
                inner_muls += 1
        muls = inner_muls + sum(1 for h in outputs if h is not None and h[0] not in (1.0, -1.0))
        return {'add': adds, 'fma': fmas, 'mul': muls,
                'flops': adds + 2 * fmas + muls, 'instructions': adds + fmas + muls}

    def evaluate(self, outputs, x, dtype=np.float64):
        """Evaluate on x (n_inputs x batch).  float32 mode rounds every node once."""
        live = self.live(outputs)
        v = {}
        for i in sorted(live):
            op = self.ops[i]
            if op[0] == 'in':
                v[i] = np.asarray(x[op[1]], dtype=dtype)
            elif op[0] == 'mul':
                v[i] = (np.float32(op[2]) * v[op[1]]).astype(np.float32) if dtype == np.float32 else op[2] * v[op[1]]
            else:
                a, b, r = v[op[1]], v[op[2]], op[3]
                if dtype == np.float32:
                    rr = np.float32(r)
                    if r == 1.0:
                        v[i] = (a + b).astype(np.float32)
                    elif r == -1.0:
                        v[i] = (a - b).astype(np.float32)
                    else:   # one rounding, as v_fma_f32
                        v[i] = (a.astype(np.float64) + np.float64(rr) * b.astype(np.float64)).astype(np.float32)
                else:
                    v[i] = a + r * b
        out = []
        for h in outputs:
            if h is None:
                out.append(np.zeros_like(np.asarray(x[0], dtype=dtype)))
            else:
                c, i = h
                if dtype == np.float32:
                    if c == 1.0:
                        out.append(v[i])
                    elif c == -1.0:
                        out.append(-v[i])
                    else:
                        out.append((np.float32(c) * v[i]).astype(np.float32))
                else:
                    out.append(c * v[i])
        return np.array(out)


# ---- complex helpers on handle pairs ------------------------------------------------

def hs(h, s):
    return None if h is None else (snap(h[0] * s), h[1])


def cadd(g, a, b):
    return (g.comb([(1.0, a[0]), (1.0, b[0])]), g.comb([(1.0, a[1]), (1.0, b[1])]))


def csub(g, a, b):
    return (g.comb([(1.0, a[0]), (-1.0, b[0])]), g.comb([(1.0, a[1]), (-1.0, b[1])]))


def cscale(a, s):
    return (hs(a[0], s), hs(a[1], s))


def cmulc(g, a, wr, wi):
    """a * (wr + i wi) for a compile-time constant."""
    return (g.comb([(wr, a[0]), (-wi, a[1])]), g.comb([(wr, a[1]), (wi, a[0])]))


def conj(a):
    return (a[0], hs(a[1], -1.0))


def kind_scale(kind, N, k):
    return {0: 1.0, 1: sj_scale(N, k), 2: sj_scale(2 * N, k), 4: sj_scale(4 * N, k)}[kind]


def cfft(g, x, kind, sigma):
    """Johnson-Frigo modified split radix on complex handles x (length N, power of two).

    Returns Y_k / s(kind, N, k), Y = sum_j x_j exp(-2 pi i sigma j k / N).
    kind 0 = plain DFT, 1 = S (/s_{N,k}), 2 = S2 (/s_{2N,k}), 4 = S4 (/s_{4N,k}).
    For real input only outputs 0..N/2 are built; the rest are conjugate mirrors.
    """
    N = len(x)
    real_in = all(v[1] is None for v in x)
    if N == 1:
        return [x[0]]
    if N == 2:
        return [cscale(cadd(g, x[0], x[1]), 1.0 / kind_scale(kind, 2, 0)),
                cscale(csub(g, x[0], x[1]), 1.0 / kind_scale(kind, 2, 1))]
    ukind = {0: 0, 1: 2, 2: 4, 4: 2}[kind]
    q = N // 4
    u = cfft(g, x[0::2], ukind, sigma)
    z = cfft(g, x[1::4], 1, sigma)
    zp = cfft(g, [x[(4 * m - 1) % N] for m in range(q)], 1, sigma)
    y = [None] * N
    for k in range(q):
        sz = sj_scale(q, k)
        ang = -2.0 * math.pi * sigma * k / N
        c, s = math.cos(ang), math.sin(ang)
        tz = cmulc(g, z[k], c * sz, s * sz)
        tzp = cmulc(g, zp[k], c * sz, -s * sz)
        A = cadd(g, tz, tzp)
        B = csub(g, tz, tzp)
        mB = (hs(B[1], sigma), hs(B[0], -sigma))          # -i sigma B
        Uk = cscale(u[k], kind_scale(ukind, N // 2, k))
        Uq = cscale(u[k + q], kind_scale(ukind, N // 2, k + q))
        y[k] = cscale(cadd(g, Uk, A), 1.0 / kind_scale(kind, N, k))
        y[k + q] = cscale(cadd(g, Uq, mB), 1.0 / kind_scale(kind, N, k + q))
        if not real_in or k == 0:
            y[k + 2 * q] = cscale(csub(g, Uk, A), 1.0 / kind_scale(kind, N, k + 2 * q))
        if not real_in:
            y[k + 3 * q] = cscale(csub(g, Uq, mB), 1.0 / kind_scale(kind, N, k + 3 * q))
    if real_in:
        for k in range(N // 2 + 1, N):
            y[k] = conj(y[N - k])
    return y


def makhoul(n):
    """v = [x0, x2, ..., x_{n-2}, x_{n-1}, ..., x3, x1] (rect_generated.inc:87 idx8_0)."""
    return [2 * j for j in range(n // 2)] + [n - 1 - 2 * j for j in range(n // 2)]


def dct2_sj(n):
    """Forward SJ core: y = F_n x with F_n[k,j] = cos(pi k (2j+1)/(2n)) / s_{4n,k}."""
    g = Graph(n)
    v = [(g.inp(j), None) for j in makhoul(n)]
    Z = cfft(g, v, 1, +1)
    out = [None] * n
    out[0] = Z[0][0]
    for k in range(1, n // 2):
        t = math.tan(math.pi * k / (2 * n))
        zr, zi = Z[k]
        out[k] = g.comb([(1.0, zr), (t, zi)])
        out[n - k] = g.comb([(t, zr), (-1.0, zi)])
    out[n // 2] = Z[n // 2][0]
    return g, out


def transpose(g, outputs, n_in):
    """Transposed linear map (transposition principle) of the DAG g with given outputs."""
    g2 = Graph(len(outputs))
    adj = {}
    for i, h in enumerate(outputs):
        if h is not None:
            adj.setdefault(h[1], []).append((h[0], g2.inp(i)))
    result = {}
    for nid in range(len(g.ops) - 1, -1, -1):
        if nid not in adj:
            continue
        a_h = g2.comb(adj.pop(nid))
        op = g.ops[nid]
        if op[0] == 'in':
            result[op[1]] = a_h
            continue
        if op[0] == 'mul':
            adj.setdefault(op[1], []).append((op[2], a_h))
            continue
        adj.setdefault(op[1], []).append((1.0, a_h))
        adj.setdefault(op[2], []).append((op[3], a_h))
    return g2, [result.get(i) for i in range(n_in)]


def dft_graph(N, sigma):
    g = Graph(2 * N)
    x = [(g.inp(j), g.inp(N + j)) for j in range(N)]
    Y = cfft(g, x, 0, sigma)
    return g, [y[0] for y in Y] + [y[1] for y in Y]


# ---- closed forms --------------------------------------------------------------------

def F_closed(n):
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    m = (k * (2 * j + 1)) % (4 * n)
    s = np.array([sj_scale(4 * n, kk) for kk in range(n)])[:, None]
    return np.cos(np.pi * m / (2.0 * n)) / s


def dft_closed_real(N, sigma):
    jk = np.outer(np.arange(N), np.arange(N)) % N
    W = np.exp(-2j * np.pi * sigma * jk / N)
    M = np.zeros((2 * N, 2 * N))
    M[:N, :N], M[:N, N:] = W.real, -W.imag
    M[N:, :N], M[N:, N:] = W.imag, W.real
    return M


def graph_matrix(g, outs, n_in):
    return g.evaluate(outs, np.eye(n_in))


def check(name, g, outs, n_in, M, tol=1e-12):
    G = graph_matrix(g, outs, n_in)
    err = np.abs(G - M).max() / np.abs(M).max()
    if not err < tol:
        raise SystemExit('%s: graph differs from closed form, rel err %.3e' % (name, err))
    return err


# ---- emission ------------------------------------------------------------------------

def flit(r):
    f = np.float32(r)
    s = '%.9e' % f
    return s + 'f'


def emit_body(g, outs, in_names, out_names):
    live = g.live(outs)
    lines = []
    name = {}
    for i in sorted(live):
        op = g.ops[i]
        t = 't%d' % i
        name[i] = t
        if op[0] == 'in':
            lines.append('    const float %s = %s;' % (t, in_names[op[1]]))
        elif op[0] == 'mul':
            lines.append('    const float %s = %s * %s;' % (t, flit(op[2]), name[op[1]]))
        else:
            a, b, r = name[op[1]], name[op[2]], op[3]
            if r == 1.0:
                e = '%s + %s' % (a, b)
            elif r == -1.0:
                e = '%s - %s' % (a, b)
            else:
                e = '__builtin_fmaf(%s, %s, %s)' % (flit(r), b, a)
            lines.append('    const float %s = %s;' % (t, e))
    for o, h in zip(out_names, outs):
        if h is None:
            lines.append('    %s = 0.0f;' % o)
            continue
        c, i = h
        if c == 1.0:
            e = name[i]
        elif c == -1.0:
            e = '-' + name[i]
        else:
            e = '%s * %s' % (flit(c), name[i])
        lines.append('    %s = %s;' % (o, e))
    return '\n'.join(lines)


def fp32_error(g, outs, n_in, M, trials=256, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n_in, trials)).astype(np.float32)
    y32 = g.evaluate(outs, x, np.float32).astype(np.float64)
    ref = M @ x.astype(np.float64)
    return float(np.abs(y32 - ref).max() / np.abs(ref).max())


KEEP_BELOW = 0.7


def bound_output(g, h):
    """Output handle (c, i) with a small pending factor c rebuilt as a node at scale c."""
    c, i = h
    if abs(c) >= KEEP_BELOW:
        return h
    op = g.ops[i]
    if op[0] == 'lin':
        return (1.0, g.lin(g.mul(op[1], c), op[2], snap(c * op[3])))
    return (1.0, g.mul(i, c))


def build_all(max_real=64, max_cplx=32, verbose=False):
    items = []   # (kind, n, sigma/inverse, g, outs, n_in, M)
    factors = {}
    n = 8
    while n <= max_real:
        g, outs = dct2_sj(n)
        F = F_closed(n)
        check('sj_%d' % n, g, outs, n, F)
        # Fold the factors still pending at the outputs into the normalization:
        # the codelet computes D^-1 F_n, D = diag(d), and F_n^T = (D^-1 F_n)^T D.
        # A pending factor below KEEP_BELOW would leave that output (the graph's largest
        # value) larger than F_n's by 1/d, which costs range near the float32 maximum;
        # such a factor is applied inside the graph instead, on the output's last node
        # (f a + (f r) b: one extra multiply), so those outputs come out at F_n's scale.
        outs = [bound_output(g, h) for h in outs]
        d = [h[0] for h in outs]
        raw = [(1.0, h[1]) for h in outs]
        M = F / np.array(d)[:, None]
        check('sjf_%d' % n, g, raw, n, M)
        items.append(('sjf', n, g, raw, n, M))
        gt, ot = transpose(g, raw, n)
        check('sjt_%d' % n, gt, ot, n, M.T)
        items.append(('sjt', n, gt, ot, n, M.T))
        factors[n] = d
        n *= 2
    N = 2
    while N <= max_cplx:
        for sigma, tag in ((+1, 'm'), (-1, 'p')):
            g, outs = dft_graph(N, sigma)
            M = dft_closed_real(N, sigma)
            check('dft_%d_%s' % (N, tag), g, outs, 2 * N, M)
            items.append(('dft_%s' % tag, N, g, outs, 2 * N, M))
        N *= 2
    return items, factors


HEADER = """// Generated by bridct_amd/tools/gen_sj.py -- do not edit.
//
// Shao-Johnson DCT-II cores and Johnson-Frigo split-radix DFT codelets for gfx950.
// sjf_n(x): x <- D_n^-1 F_n x, F_n[k][j] = cos(pi k (2j+1) / (2n)) / s_{4n,k} (reference
//           src/plan.c:39-47, :96), D_n = diag(bridct_sj_factor_n) from sj_factors.h;
// sjt_n(x): x <- (D_n^-1 F_n)^T x.
// dft_N_m(re, im): X_k = sum_j x_j exp(-2 pi i j k / N); dft_N_p: exp(+2 pi i j k / N).
// Every node is one add, subtract or FMA (single rounding); the graphs were checked in
// FP64 against the closed forms when this file was written.
#pragma once
"""


def write_factors(factors, outdir):
    """Host table: d_n[k] with sjf_n = diag(1/d_n) F_n (FP64, for the plan's tables)."""
    lines = ['/* Generated by bridct_amd/tools/gen_sj.py -- do not edit.',
             ' * sjf_n computes diag(1/d_n) F_n; the plan multiplies d_n into its FP64',
             ' * normalization before the single rounding to FP32. */',
             '#ifndef BRIDCT_SJ_FACTORS_H', '#define BRIDCT_SJ_FACTORS_H']
    for n in sorted(factors):
        vals = ', '.join('%.17g' % v for v in factors[n])
        lines.append('static const double bridct_sj_factor_%d[%d] = {%s};' % (n, n, vals))
    lines.append('#endif')
    with open(os.path.join(outdir, 'sj_factors.h'), 'w') as f:
        f.write('\n'.join(lines) + '\n')


def write_outputs(items, outdir):
    os.makedirs(outdir, exist_ok=True)
    parts = [HEADER]
    counts = {}
    for kind, n, g, outs, n_in, M in items:
        c = g.counts(outs)
        c['fp32_max_rel_err'] = fp32_error(g, outs, n_in, M)
        if kind in ('sjf', 'sjt'):
            fname = '%s_%d' % (kind, n)
            names = ['x[%d]' % i for i in range(n)]
            sig = 'static __device__ __forceinline__ void %s(float (&x)[%d])' % (fname, n)
            body = emit_body(g, outs, names, names)
        else:
            tag = kind[-1]
            fname = 'dft_%d_%s' % (n, tag)
            names = ['re[%d]' % i for i in range(n)] + ['im[%d]' % i for i in range(n)]
            sig = ('static __device__ __forceinline__ void %s(float (&re)[%d], float (&im)[%d])'
                   % (fname, n, n))
            body = emit_body(g, outs, names, names)
        counts[fname] = c
        parts.append('// %s: %d add/sub, %d fma, %d mul\n%s\n{\n%s\n}\n'
                     % (fname, c['add'], c['fma'], c['mul'], sig, body))
    with open(os.path.join(outdir, 'sj_codelets.h'), 'w') as f:
        f.write('\n'.join(parts))
    with open(os.path.join(outdir, 'sj_counts.json'), 'w') as f:
        json.dump(counts, f, indent=1, sort_keys=True)
    return counts


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--out', default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  '..', 'csrc', 'generated'))
    ap.add_argument('--check', action='store_true', help='check graphs only, write nothing')
    ap.add_argument('--counts-to', type=int, default=0,
                    help='also check and count F_n up to this n (no emission)')
    args = ap.parse_args(argv)
    items, factors = build_all()
    if args.counts_to:
        n = 128
        while n <= args.counts_to:
            g, outs = dct2_sj(n)
            check('sjf_%d' % n, g, outs, n, F_closed(n), tol=1e-11)
            print('sjf_%d' % n, g.counts(outs))
            n *= 2
    if args.check:
        for kind, n, g, outs, n_in, M in items:
            print(kind, n, g.counts(outs))
        return 0
    counts = write_outputs(items, os.path.normpath(args.out))
    write_factors(factors, os.path.normpath(args.out))
    for k in sorted(counts):
        print(k, counts[k])
    return 0


if __name__ == '__main__':
    sys.exit(main())

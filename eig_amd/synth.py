"""Seeded synthetic genotype panels (Balding-Nichols), identical on CPU and GPU.

Population allele frequencies are drawn on the host with numpy's PCG64:
  ancestral p ~ U(plo, phi); population k: p_k ~ Beta(p(1-F_k)/F_k, (1-p)(1-F_k)/F_k).
They become uint32 thresholds thr[k][m] = floor(p_k * 2^32).  Genotype bits come
from a counter-based hash of (seed, m*N + i): allele copy c of sample i picks its
source population from a cumulative-weight draw (admixture), and is the first
allele when its uniform falls below that population's threshold; a third draw
below miss_thr marks the genotype missing.  Everything after the thresholds is
integer arithmetic, so eig_synth_packed on the device and `genotypes()` here give
the same bits.
"""
from __future__ import annotations

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(x):
    x = x ^ (x >> np.uint64(30))
    x = x * np.uint64(0xBF58476D1CE4E5B9)
    x = x ^ (x >> np.uint64(27))
    x = x * np.uint64(0x94D049BB133111EB)
    x = x ^ (x >> np.uint64(31))
    return x


class Panel:
    """Parameters of one synthetic panel (what eig_synth_packed consumes)."""

    def __init__(self, nsnp, pop_sizes, fst, plo, phi, miss, seed, admix_frac=0.0,
                 admix_k=3):
        rng = np.random.Generator(np.random.PCG64(seed))
        self.nsnp = int(nsnp)
        self.npop = len(pop_sizes)
        self.nind = int(sum(pop_sizes))
        self.seed = int(seed)
        fst = np.broadcast_to(np.asarray(fst, dtype=np.float64), (self.npop,))
        p = rng.uniform(plo, phi, size=self.nsnp)
        pk = np.empty((self.npop, self.nsnp))
        for k in range(self.npop):
            f = fst[k]
            a = p * (1 - f) / f
            b = (1 - p) * (1 - f) / f
            pk[k] = rng.beta(a, b)
        self.pk = pk
        self.thr = np.minimum(np.floor(pk * 4294967296.0), 4294967295.0).astype(np.uint32)
        pop_of = np.repeat(np.arange(self.npop, dtype=np.int32), pop_sizes)
        self.pop_of = pop_of
        anc = np.repeat(pop_of[:, None], 3, axis=1).astype(np.int32)
        cum = np.full((self.nind, 2), 0xFFFFFFFF, dtype=np.uint32)
        self.admixed = np.zeros(self.nind, dtype=bool)
        if admix_frac > 0:
            nad = int(round(admix_frac * self.nind))
            who = rng.choice(self.nind, size=nad, replace=False)
            for i in who:
                pops = rng.choice(self.npop, size=admix_k, replace=False)
                w = rng.dirichlet(np.ones(admix_k))
                c1 = min(int(w[0] * 4294967296.0), 4294967295)
                c2 = min(int((w[0] + w[1]) * 4294967296.0), 4294967295)
                anc[i, :] = pops[:3] if admix_k >= 3 else np.resize(pops, 3)
                cum[i] = (c1, c2)
                self.admixed[i] = True
        self.anc = anc
        self.cum = cum
        self.miss_thr = np.uint32(min(int(miss * 4294967296.0), 4294967295))

    # ---- CPU restatement of k_synth (small sizes)
    def genotypes(self, m0=0, m1=None):
        """int32 [m1-m0, nind] codes 0/1/2, 3 = missing."""
        m1 = self.nsnp if m1 is None else m1
        m = np.arange(m0, m1, dtype=np.uint64)[:, None]
        i = np.arange(self.nind, dtype=np.uint64)[None, :]
        key = m * np.uint64(self.nind) + i
        with np.errstate(over="ignore"):
            h1 = _mix64(np.uint64(self.seed) + (key + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
            h2 = _mix64(h1 ^ np.uint64(0x5851F42D4C957F2D))
            h3 = _mix64(h2 + np.uint64(0x2545F4914F6CDD1D))
        g = np.zeros(key.shape, dtype=np.int32)
        mi = np.arange(m0, m1)[:, None]
        for c in range(2):
            ua = (h1 >> np.uint64(32 * c)) & np.uint64(0xFFFFFFFF)
            up = (h2 >> np.uint64(32 * c)) & np.uint64(0xFFFFFFFF)
            c0 = self.cum[:, 0].astype(np.uint64)[None, :]
            c1 = self.cum[:, 1].astype(np.uint64)[None, :]
            pop = np.where(up < c0, self.anc[:, 0][None, :],
                           np.where(up < c1, self.anc[:, 1][None, :], self.anc[:, 2][None, :]))
            thr = self.thr[pop, mi].astype(np.uint64)
            g += (ua < thr).astype(np.int32)
        miss = (h3 & np.uint64(0xFFFFFFFF)) < np.uint64(self.miss_thr)
        g[miss] = 3
        return g

    def packed(self, m0=0, m1=None):
        g = self.genotypes(m0, m1)
        M, N = g.shape
        rlen = max(48, (2 * N + 7) // 8)
        pad = np.zeros((M, rlen * 4), np.uint8)
        pad[:, :N] = g
        pad = pad.reshape(M, rlen, 4)
        return (pad[:, :, 0] << 6 | pad[:, :, 1] << 4 | pad[:, :, 2] << 2 | pad[:, :, 3]).astype(np.uint8)

    def to_device(self, eng):
        eng.synth(self.nsnp, self.nind, self.thr, self.anc, self.cum, self.miss_thr, self.seed)

    def ids(self):
        inds = ["S%06d" % i for i in range(self.nind)]
        pops = ["P%02d" % (k + 1) for k in self.pop_of]
        snps = ["rs%07d" % m for m in range(self.nsnp)]
        return inds, pops, snps


def query_genotypes(thr, pop, miss_thr, seed, nsnp=None):
    """CPU restatement of eig_synth_queries: int32 [nq, nsnp] codes, 3 = missing.
    Sample q draws both allele copies from population pop[q]; key = q*nsnp + m."""
    thr = np.asarray(thr, dtype=np.uint32)
    nsnp = thr.shape[1] if nsnp is None else nsnp
    pop = np.asarray(pop)
    q = np.arange(len(pop), dtype=np.uint64)[:, None]
    m = np.arange(nsnp, dtype=np.uint64)[None, :]
    key = q * np.uint64(nsnp) + m
    with np.errstate(over="ignore"):
        h1 = _mix64(np.uint64(seed) + (key + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
        h2 = _mix64(h1 ^ np.uint64(0x5851F42D4C957F2D))
    t = thr[pop][:, :nsnp].astype(np.uint64)
    lo = h1 & np.uint64(0xFFFFFFFF)
    hi = h1 >> np.uint64(32)
    g = (lo < t).astype(np.int32) + (hi < t).astype(np.int32)
    miss = (h2 & np.uint64(0xFFFFFFFF)) < np.asarray(miss_thr, dtype=np.uint64)[:, None]
    g[miss] = 3
    return g


def pack_samples(g: np.ndarray, stride=None) -> np.ndarray:
    """Sample-major packed rows of codes g [nq, nsnp]: SNP m at byte m/4, bits
    7-6, 5-4, 3-2, 1-0 for m%4 = 0..3; padding codes are 3."""
    nq, M = g.shape
    nb = (M + 3) // 4
    stride = nb if stride is None else stride
    pad = np.full((nq, stride * 4), 3, np.uint8)
    pad[:, :M] = g
    pad = pad.reshape(nq, stride, 4)
    return (pad[:, :, 0] << 6 | pad[:, :, 1] << 4 | pad[:, :, 2] << 2 | pad[:, :, 3]).astype(np.uint8)


def serving_queries(P: Panel, nq: int, seed: int, miss_lo=0.0, miss_hi=0.9):
    """Populations and per-sample missingness thresholds of a query set drawn from
    the panel's populations (missing fraction ~ U(miss_lo, miss_hi))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pop = rng.integers(0, P.npop, size=nq).astype(np.int32)
    frac = rng.uniform(miss_lo, miss_hi, size=nq)
    miss_thr = np.minimum(np.floor(frac * 4294967296.0), 4294967295.0).astype(np.uint32)
    return pop, miss_thr


def config_panel(cfg: int, nsnp=None, scale=1.0):
    """The five workload shapes (index 1..5); nsnp / scale shrink them for tests."""
    seed = 0xE16 + cfg
    if cfg == 1:
        sizes = [334, 333, 333]
        P = Panel(nsnp or 50000, [int(s * scale) for s in sizes], 0.05, 0.05, 0.5, 0.01, seed)
    elif cfg in (2, 5):
        rng = np.random.Generator(np.random.PCG64(seed ^ 0x55))
        fst = rng.uniform(0.01, 0.1, size=10)
        P = Panel(nsnp or 600000, [int(500 * scale)] * 10, fst, 0.01, 0.5, 0.02, 0xE16 + 2)
    elif cfg == 3:
        rng = np.random.Generator(np.random.PCG64(seed ^ 0x55))
        fst = rng.uniform(0.005, 0.08, size=25)
        P = Panel(nsnp or 1000000, [int(800 * scale)] * 25, fst, 0.01, 0.5, 0.03, seed,
                  admix_frac=0.05)
    elif cfg == 4:
        rng = np.random.Generator(np.random.PCG64(seed ^ 0x55))
        fst = rng.uniform(0.005, 0.05, size=50)
        P = Panel(nsnp or 500000, [int(2000 * scale)] * 50, fst, 0.01, 0.5, 0.01, seed)
    else:
        raise ValueError(cfg)
    return P


def hasharr(ids) -> int:
    """GENO header hash of an ID list (admutils.c:651-682), 32-bit wraparound."""
    h = 0
    for s in ids:
        t = 0
        for ch in s.encode():
            t = (t * 23 + (ch if ch < 128 else ch - 256)) & 0xFFFFFFFF
        h = ((h * 17) & 0xFFFFFFFF) ^ t
    return h


def write_files(P: Panel, prefix: str):
    """Writes prefix.geno (packed GENO with header hash), .snp, .ind."""
    hash_fn = hasharr
    inds, pops, snps = P.ids()
    with open(prefix + ".ind", "w") as f:
        for i, p in zip(inds, pops):
            f.write("%20s U %10s\n" % (i, p))
    with open(prefix + ".snp", "w") as f:
        for m, s in enumerate(snps):
            chrom = 1 + (m * 22) // P.nsnp
            f.write("%20s %3d %12.6f %12d A G\n" % (s, chrom, 0.0, 1000 + 100 * m))
    rlen = max(48, (2 * P.nind + 7) // 8)
    hdr = ("GENO %7d %7d %x %x" % (P.nind, P.nsnp, hash_fn(inds) & 0xFFFFFFFF,
                                    hash_fn(snps) & 0xFFFFFFFF)).encode().ljust(rlen, b"\0")
    with open(prefix + ".geno", "wb") as f:
        f.write(hdr)
        step = max(1, 2000000 // max(1, P.nind))
        for m0 in range(0, P.nsnp, step):
            f.write(P.packed(m0, min(P.nsnp, m0 + step)).tobytes())

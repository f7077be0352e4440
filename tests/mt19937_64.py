"""Pure-Python std::mt19937_64 (Matsumoto & Nishimura 64-bit Mersenne
Twister, parameters of C++11 [rand.predef]) and the forcing draw that
libsipg_hip's sipg_forcing_draw documents.  Test helper: it pins the
library's draw without running the library."""

_MASK = (1 << 64) - 1


class MT19937_64:
    n, m = 312, 156
    a = 0xB5026F5AA96619E9
    upper, lower = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed=5489):
        self.mt = [0] * self.n
        self.mt[0] = seed & _MASK
        for i in range(1, self.n):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & _MASK
        self.idx = self.n

    def _twist(self):
        mt = self.mt
        for i in range(self.n):
            x = (mt[i] & self.upper) | (mt[(i + 1) % self.n] & self.lower)
            xa = x >> 1
            if x & 1:
                xa ^= self.a
            mt[i] = mt[(i + self.m) % self.n] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= self.n:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK


def forcing_draw(seed, kmin, kmax, mmax):
    g = MT19937_64(seed)
    K = kmin + g() % (kmax - kmin + 1)
    a, m, n = [], [], []
    for _ in range(K):
        a.append(-1.0 + 2.0 * ((g() >> 11) * 2.0 ** -53))
        m.append(1 + g() % mmax)
        n.append(1 + g() % mmax)
    return a, m, n

"""Pure-Python big-int restatements used to pin the oracle on small cases.

These are textbook definitions (no Montgomery, no windowing): modular
arithmetic with Python ints, affine elliptic-curve addition over Fp and
Fp2 = Fp[u]/(u^2+1), DFT by definition and polynomial evaluation by Horner.
"""
P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
RM = 1 << 256

G1 = (1, 2)
G2 = (
    (0x1800DEEF121F1E76426A00665E5C4479674322D4F75EDADD46DEBD5CD992F6ED,
     0x198E9393920D483A7260BFB731FB5D25F1AA493335A9E71297E485B7AEF312C2),
    (0x12C85EA5DB8C6DEB4AAB71808DCB408FE3D1E7690C43D37B4CE6CC0166FA7DAA,
     0x090689D0585FF075EC9E99AD690C3395BC4B313370B38EF355ACDADCD122975B),
)


def limbs(x, n=4):
    return [(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(n)]


def from_limbs(ls):
    v = 0
    for i, w in enumerate(ls):
        v |= int(w) << (64 * i)
    return v


def root(log_n):
    return pow(5, (R - 1) >> log_n, R)


def dft(xs, w):
    n = len(xs)
    return [sum(xs[j] * pow(w, j * k, R) for j in range(n)) % R for k in range(n)]


def interp_eval(evals, x):
    """Evaluate at x the degree < n interpolant of evals on the size-n domain."""
    n = len(evals)
    log_n = n.bit_length() - 1
    w = root(log_n)
    winv = pow(w, R - 2, R)
    ninv = pow(n, R - 2, R)
    coeffs = [sum(evals[k] * pow(winv, j * k, R) for k in range(n)) * ninv % R for j in range(n)]
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % R
    return acc


# ---- Fp2 and affine curve arithmetic (None = point at infinity) ----
def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2_inv(a):
    n = pow(a[0] * a[0] + a[1] * a[1], P - 2, P)
    return (a[0] * n % P, -a[1] * n % P)


class Fp1:
    @staticmethod
    def add(a, b): return (a + b) % P
    @staticmethod
    def sub(a, b): return (a - b) % P
    @staticmethod
    def mul(a, b): return a * b % P
    @staticmethod
    def inv(a): return pow(a, P - 2, P)
    @staticmethod
    def small(k): return k % P
    zero = 0


class Fp2:
    add = staticmethod(f2_add)
    sub = staticmethod(f2_sub)
    mul = staticmethod(f2_mul)
    inv = staticmethod(f2_inv)
    @staticmethod
    def small(k): return (k % P, 0)
    zero = (0, 0)


def ec_add(p, q, F):
    if p is None:
        return q
    if q is None:
        return p
    if p[0] == q[0]:
        if F.add(p[1], q[1]) == F.zero:
            return None
        lam = F.mul(F.mul(F.small(3), F.mul(p[0], p[0])), F.inv(F.mul(F.small(2), p[1])))
    else:
        lam = F.mul(F.sub(q[1], p[1]), F.inv(F.sub(q[0], p[0])))
    x3 = F.sub(F.sub(F.mul(lam, lam), p[0]), q[0])
    y3 = F.sub(F.mul(lam, F.sub(p[0], x3)), p[1])
    return (x3, y3)


def ec_mul(p, k, F):
    acc = None
    for bit in bin(k)[2:] if k else "":
        acc = ec_add(acc, acc, F)
        if bit == "1":
            acc = ec_add(acc, p, F)
    return acc

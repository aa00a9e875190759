"""Textbook secp256k1 arithmetic in Python integers, used only as an
independent mathematical check of the device primitives (closed forms:
modular inverse via pow, affine group law, GLV identity k = k1 + lambda*k2)."""
P = 2 ** 256 - 2 ** 32 - 977
N = 0xFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFFEBAAEDCE6AF48A03BBFD25E8CD0364141
GX = 0x79BE667EF9DCBBAC55A06295CE870B07029BFCDB2DCE28D959F2815B16F81798
GY = 0x483ADA7726A3C4655DA4FBFC0E1108A8FD17B448A68554199C47D08FFB10D4B8
G = (GX, GY)
LAMBDA = 0x5363AD4CC05C30E0A5261C028812645A122E22EA20816678DF02967C1B23BD72
BETA = 0x7AE96A2B657C07106E64479EAC3434E99CF0497512F58995C1396C28719501EE


def add(p1, p2):
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return (x3, (lam * (x1 - x3) - y1) % P)


def mul(k, pt=G):
    r = None
    for b in bin(k % N)[2:] if k % N else "":
        r = add(r, r)
        if b == "1":
            r = add(r, pt)
    return r


def lift_x(x, odd=False):
    if x >= P:
        return None
    y2 = (pow(x, 3, P) + 7) % P
    y = pow(y2, (P + 1) // 4, P)
    if y * y % P != y2:
        return None
    if (y & 1) != int(odd):
        y = P - y
    return (x, y)


def compress(pt):
    return bytes([2 | (pt[1] & 1)]) + pt[0].to_bytes(32, "big")


def to_limbs(v, n=8):
    return [(v >> (32 * i)) & 0xFFFFFFFF for i in range(n)]


def from_limbs(ws):
    return sum(int(w) << (32 * i) for i, w in enumerate(ws))

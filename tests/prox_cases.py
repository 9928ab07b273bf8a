"""Known-answer cases of the PROX contact reduction (prox_collision_detection.h:
340-345: std::sort, then std::unique over prox::ContactPoint, prox_contact_point.h
:91-190), each with its expected output worked out by hand from the reference's
rules, not by running a restatement of them:

- order: body i, body j, then p.y, p.x, p.z, n.y, n.x, n.z, depth;
- equal: same bodies and every one of those floats within 10 * FLT_EPSILON =
  1.1920929e-06 (|difference| > it is unequal, so a difference of exactly that
  much is still equal);
- std::unique (libstdc++) keeps a contact unless it equals the last contact
  KEPT, not the one just before it.

At 1.0 one float ulp is 2^-23, so 10 ulps there is exactly the precision.
Each case: (name, records, kept), kept = input positions of the survivors in
output order.
"""
import numpy as np

DTYPE = np.dtype([("body_i", np.uint32), ("body_j", np.uint32), ("p", np.float32, (3,)), ("n", np.float32, (3,)),
                  ("depth", np.float32)])

ULP1 = np.float32(2.0 ** -23)  # one ulp of floats in [1, 2)


def _rec(rows):
    a = np.zeros(len(rows), DTYPE)
    for i, (bi, bj, p, n, d) in enumerate(rows):
        a[i] = (bi, bj, p, n, d)
    return a


def _y(k):
    """1.0 + k ulps, as float32."""
    return np.float32(1.0) + np.float32(k) * ULP1


N0 = (0.0, 1.0, 0.0)


def cases():
    out = []
    # Exact duplicates: the first of each group survives, order by p.y.
    out.append(("exact_duplicates", _rec([
        (0, 1, (0.0, 2.0, 0.0), N0, -0.1),
        (0, 1, (0.0, 1.0, 0.0), N0, -0.1),
        (0, 1, (0.0, 2.0, 0.0), N0, -0.1),
        (0, 1, (0.0, 1.0, 0.0), N0, -0.1),
    ]), [1, 0]))
    # p.y is compared before p.x: (x=5, y=0) sorts before (x=0, y=1).
    out.append(("py_before_px", _rec([
        (0, 1, (0.0, 1.0, 0.0), N0, -0.1),
        (0, 1, (5.0, 0.0, 0.0), N0, -0.1),
    ]), [1, 0]))
    # Bodies first: equal geometry under other bodies is never merged; body j
    # orders within body i.
    out.append(("bodies_first", _rec([
        (2, 0, (0.0, 0.0, 0.0), N0, 0.0),
        (0, 3, (0.0, 0.0, 0.0), N0, 0.0),
        (0, 1, (0.0, 0.0, 0.0), N0, 0.0),
        (0, 1, (0.0, 0.0, 0.0), N0, 0.0),
    ]), [2, 1, 0]))
    # The precision boundary: 10 ulps at 1.0 equals, 11 ulps does not.
    out.append(("precision_boundary", _rec([
        (0, 1, (0.0, _y(0), 0.0), N0, 0.0),
        (0, 1, (0.0, _y(10), 0.0), N0, 0.0),
        (1, 1, (0.0, _y(0), 0.0), N0, 0.0),
        (1, 1, (0.0, _y(11), 0.0), N0, 0.0),
    ]), [0, 2, 3]))
    # The chain rule: y, y+10u, y+11u.  The second equals the first (dropped);
    # the third is 1 ulp from the second but 11 from the last kept (the first),
    # so it is kept.  Comparing with the previous contact would drop it.
    out.append(("unique_compares_with_last_kept", _rec([
        (0, 1, (0.0, _y(11), 0.0), N0, 0.0),
        (0, 1, (0.0, _y(0), 0.0), N0, 0.0),
        (0, 1, (0.0, _y(10), 0.0), N0, 0.0),
    ]), [1, 0]))
    # A longer chain: 0, 6, 12, 18, 24 ulps -> kept 0 (6 dropped), 12 (18
    # dropped), 24.
    out.append(("chain_of_five", _rec([(0, 1, (0.0, _y(k), 0.0), N0, 0.0) for k in (24, 18, 12, 6, 0)]), [4, 2, 0]))
    # Every field takes part in equality: normal x 11 ulps apart (at 1.0) is
    # unequal, depth 5 ulps apart is equal.
    out.append(("normal_and_depth", _rec([
        (0, 1, (0.0, 0.0, 0.0), (_y(0), 0.0, 0.0), _y(0)),
        (0, 1, (0.0, 0.0, 0.0), (_y(11), 0.0, 0.0), _y(0)),
        (0, 1, (0.0, 0.0, 0.0), (_y(0), 0.0, 0.0), _y(5)),
    ]), [0, 1]))
    # Sorted by p.y: (y0, x=3), (y0+4u, x=0), (y0+8u, x=3).  The third equals
    # the first (p.y 8 ulps apart, x equal), but the second was kept in
    # between (x 3 apart), so the third is compared with the second and kept:
    # no merging with an earlier kept contact.
    out.append(("later_key_breaks_equality", _rec([
        (0, 1, (3.0, _y(8), 0.0), N0, 0.0),
        (0, 1, (0.0, _y(4), 0.0), N0, 0.0),
        (0, 1, (3.0, _y(0), 0.0), N0, 0.0),
    ]), [2, 1, 0]))
    # Signed zeros compare equal under both operators; the first in input
    # order survives.
    out.append(("signed_zero", _rec([
        (0, 1, (-0.0, 0.0, 0.0), N0, 0.0),
        (0, 1, (0.0, 0.0, 0.0), N0, 0.0),
    ]), [0]))
    out.append(("single", _rec([(7, 7, (1.0, 2.0, 3.0), N0, -1.0)]), [0]))
    return out

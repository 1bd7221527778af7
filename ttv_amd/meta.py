"""Output metadata helpers for the Python front end (same rules as include/tlib/detail):

* generate_output_shape  -- the extents of A without mode q             (shape.h:103-158)
* generate_output_layout -- A's layout without q, modes above q renumbered (layout.h:143-207)
* generate_strides       -- dense strides, all ones for vector/scalar shapes (strides.h:31-74)
* generate_k_order_layout -- {k,...,1,k+1,...,p}, k = 0 for last-order   (layout.h:57-84)
* layout_from_strides    -- the layout tuple whose strides are non-decreasing (ties: last-order)
"""


def generate_output_shape(na, q):
    return [n for r, n in enumerate(na, 1) if r != q]


def generate_output_layout(pia, q):
    return [m - 1 if m > q else m for m in pia if m != q]


def _is_vector_or_scalar(n):
    if not n or any(x == 0 for x in n):
        return False
    if all(x == 1 for x in n):
        return True
    if len(n) == 1:
        return n[0] > 1
    return (n[0] > 1 or n[1] > 1) and (n[0] == 1 or n[1] == 1) and all(x == 1 for x in n[2:])


def generate_strides(n, pi):
    p = len(n)
    w = [1] * p
    if _is_vector_or_scalar(n):
        return w
    for r in range(1, p):
        w[pi[r] - 1] = w[pi[r - 1] - 1] * n[pi[r - 1] - 1]
    return w


def generate_k_order_layout(p, k):
    if k == 0:
        k = p
    m = min(k, p)
    return list(range(m, 0, -1)) + list(range(m + 1, p + 1))


def layout_from_strides(n, w):
    """Modes ordered by increasing stride; equal strides keep the last-order convention
    (higher mode first), which is what a C-contiguous array with unit extents needs."""
    p = len(n)
    return sorted(range(1, p + 1), key=lambda m: (w[m - 1], -m))

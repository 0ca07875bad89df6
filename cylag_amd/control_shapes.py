"""Control sections: lines across the flow through which the solver counts
particle crossings at every step (cylag/extra/control_shapes.pyx:7-85).

The counting runs on the device inside the step (k_control_sections in
csrc/cyl_kernels.h); this class only carries the polygon.  As in the
reference, the number of segments tested is ``poly.shape[1] - 1``: for an
(npoint, 2) polygon that is its first segment [poly[0], poly[1]] only.
"""
import numpy as np


class ControlSection:
    def __init__(self, poly):
        self.poly = np.ascontiguousarray(poly, dtype='d')
        if self.poly.ndim != 2:
            raise ValueError("ControlSection: poly must be a (npoint, 2) array")
        self.nseg = self.poly.shape[1] - 1

    def segments(self):
        """(nseg, 4) array of the tested segments (x0, y0, x1, y1)."""
        q = self.poly
        return np.array([[q[k, 0], q[k, 1], q[k + 1, 0], q[k + 1, 1]] for k in range(self.nseg)],
                        dtype='d').reshape(-1, 4)


def section_segments(control_sections):
    """(nseg_of[nsec], seg[nseg_total * 4]) for cyl_control_sections_set."""
    segs = [cs.segments() for cs in control_sections]
    nseg_of = np.array([s.shape[0] for s in segs], dtype=np.int32)
    seg = np.ascontiguousarray(np.concatenate(segs).reshape(-1) if segs else np.zeros(0), dtype='d')
    return nseg_of, seg

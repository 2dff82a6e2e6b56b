"""A base view with rectified partners in numpy, after the rig of the
reference's tests/test_mvs.cpp:23-120 (cameras, slanted plane, rotations):
the rectified pairs' geometry (PairGeometry of mvs/multi_stereo.cpp:26-33) and
exact plane disparities, computed here independently of the library."""
import math

import numpy as np

from regard3d_mvs_amd import GridCamera, PairGeometry


def rotation(rx_deg, ry_deg, rz_deg):
    """Rz * Ry * Rx (world to camera rows), as test_mvs.cpp:25-31."""
    k = math.pi / 180.0
    a, b, c = rx_deg * k, ry_deg * k, rz_deg * k
    rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    rz = np.array([[math.cos(c), -math.sin(c), 0], [math.sin(c), math.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


W, H, F, CX, CY = 320, 240, 280.0, 161.2, 118.7


def kmat(f, cx, cy):
    return np.array([[f, 0, cx], [0, f, cy], [0, 0, 1.0]])


class Rig:
    def __init__(self, offsets=((0.8, 0.05, 0.1), (-0.7, 0.1, 0.0), (0.1, 0.75, -0.1))):
        self.rb = rotation(2.0, -3.0, 1.0)
        self.cb = np.array([0.1, -0.2, 0.0])
        self.partners = [self.cb + self.rb.T @ np.array(o) for o in offsets]
        view = self.rb[2]
        n = view + self.rb.T @ np.array([0.3, -0.2, 0.0])
        self.normal = n / np.linalg.norm(n)
        self.point = self.cb + 5.0 * view

    def grid(self):
        return GridCamera(W, H, tuple(np.linalg.inv(kmat(F, CX, CY)).ravel()), tuple(self.rb.T.ravel()),
                          tuple(self.cb))

    def depth(self, x, y):
        """Exact plane depth of grid pixel (x, y) (distance along the optical axis)."""
        ray = self.rb.T @ np.linalg.inv(kmat(F, CX, CY)) @ np.array([x, y, 1.0])
        return (self.point - self.cb) @ self.normal / (ray @ self.normal)

    def pair(self, cp, shift=0.0, confidence=0.5):
        """Rectified pair of the base view with partner centre cp: x axis along
        the baseline, K_r = (F, CX, CY), partner cx 20 px smaller."""
        e1 = (cp - self.cb) / np.linalg.norm(cp - self.cb)
        e2 = np.cross(self.rb[2], e1)
        e2 /= np.linalg.norm(e2)
        e3 = np.cross(e1, e2)
        rr = np.stack([e1, e2, e3])
        kl, kr = kmat(F, CX, CY), kmat(F, CX - 20.0, CY)
        ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
        pix = np.stack([xs, ys, np.ones_like(xs)], -1)
        rays = pix @ (rr.T @ np.linalg.inv(kl)).T
        t = ((self.point - self.cb) @ self.normal) / (rays @ self.normal)
        xw = self.cb + t[..., None] * rays
        pc = (xw - cp) @ rr.T
        xr = kr[0, 0] * pc[..., 0] / pc[..., 2] + kr[0, 2]
        disp = (xs - xr + shift).astype(np.float32)
        g = kl @ rr @ self.rb.T @ np.linalg.inv(kmat(F, CX, CY))
        return PairGeometry(tuple(g.ravel()), F * np.linalg.norm(cp - self.cb), 20.0, tuple(cp), disp,
                            np.full((H, W), confidence, np.float32))

"""Synthetic multi-view scenes for combine_pairs: a base view with rectified
partners in numpy, after the rig of the reference's tests/test_mvs.cpp:23-120
(cameras, slanted plane, rotations).  Gives the rectified pairs' geometry
(PairGeometry of mvs/multi_stereo.cpp:26-33), exact plane disparities and the
exact plane depth of every grid pixel; plain numpy, independent of the GPU
library and of the CPU oracle."""
import math

import numpy as np

from ._abi import GridCamera, PairGeometry


def rotation(rx_deg, ry_deg, rz_deg):
    """Rz * Ry * Rx (world to camera rows), as test_mvs.cpp:25-31."""
    k = math.pi / 180.0
    a, b, c = rx_deg * k, ry_deg * k, rz_deg * k
    rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    rz = np.array([[math.cos(c), -math.sin(c), 0], [math.sin(c), math.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


def kmat(f, cx, cy):
    return np.array([[f, 0, cx], [0, f, cy], [0, 0, 1.0]])


class Rig:
    """scale: image size and focal length relative to the reference's 320x240
    camera with f = 280."""

    def __init__(self, offsets=((0.8, 0.05, 0.1), (-0.7, 0.1, 0.0), (0.1, 0.75, -0.1)), scale=1.0):
        self.W, self.H = int(round(320 * scale)), int(round(240 * scale))
        self.F, self.CX, self.CY = 280.0 * scale, 161.2 * scale, 118.7 * scale
        self.rb = rotation(2.0, -3.0, 1.0)
        self.cb = np.array([0.1, -0.2, 0.0])
        self.partners = [self.cb + self.rb.T @ np.array(o) for o in offsets]
        view = self.rb[2]
        n = view + self.rb.T @ np.array([0.3, -0.2, 0.0])
        self.normal = n / np.linalg.norm(n)
        self.point = self.cb + 5.0 * view

    def grid(self):
        return GridCamera(self.W, self.H, tuple(np.linalg.inv(kmat(self.F, self.CX, self.CY)).ravel()), tuple(self.rb.T.ravel()),
                          tuple(self.cb))

    def depth(self, x, y):
        """Exact plane depth of grid pixel (x, y) (distance along the optical axis)."""
        ray = self.rb.T @ np.linalg.inv(kmat(self.F, self.CX, self.CY)) @ np.array([x, y, 1.0])
        return (self.point - self.cb) @ self.normal / (ray @ self.normal)

    def pair(self, cp, shift=0.0, confidence=0.5):
        """Rectified pair of the base view with partner centre cp: x axis along
        the baseline, K_r = (f, cx, cy) of the rig, partner cx 20 px smaller."""
        e1 = (cp - self.cb) / np.linalg.norm(cp - self.cb)
        e2 = np.cross(self.rb[2], e1)
        e2 /= np.linalg.norm(e2)
        e3 = np.cross(e1, e2)
        rr = np.stack([e1, e2, e3])
        kl, kr = kmat(self.F, self.CX, self.CY), kmat(self.F, self.CX - 20.0, self.CY)
        ys, xs = np.mgrid[0:self.H, 0:self.W].astype(np.float64)
        m = rr.T @ np.linalg.inv(kl)  # rectified left pixel -> world ray
        r = [m[i, 0] * xs + m[i, 1] * ys + m[i, 2] for i in range(3)]
        t = ((self.point - self.cb) @ self.normal) / (r[0] * self.normal[0] + r[1] * self.normal[1] +
                                                       r[2] * self.normal[2])
        q = [self.cb[i] - cp[i] + t * r[i] for i in range(3)]  # world point minus the partner centre
        px = rr[0, 0] * q[0] + rr[0, 1] * q[1] + rr[0, 2] * q[2]
        pz = rr[2, 0] * q[0] + rr[2, 1] * q[1] + rr[2, 2] * q[2]
        xr = kr[0, 0] * px / pz + kr[0, 2]
        disp = (xs - xr + shift).astype(np.float32)
        g = kl @ rr @ self.rb.T @ np.linalg.inv(kmat(self.F, self.CX, self.CY))
        return PairGeometry(tuple(g.ravel()), self.F * np.linalg.norm(cp - self.cb), 20.0, tuple(cp), disp,
                            np.full((self.H, self.W), confidence, np.float32))

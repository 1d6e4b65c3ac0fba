"""Uniform random rotations and quaternion -> matrix conversion.

Restates ``emsim/utils/rot.py`` (reference): Shoemake's uniform quaternion
draw (``:8-36``, three ``np.random.random`` numbers per quaternion, so a
seeded global RNG reproduces the reference's draws), the batched draw
(``:39-53``), the axis-angle quaternion (``:82-104``) and the matrix
conversion (``:107-157``).  The reference's axis-angle matrix helper
(``:56-79``) has no caller on the hot path and is not restated.  The matrix keeps the reference's convention,
which is the transpose of the usual active rotation matrix
(``rot[0, 1] = 2 (q1 q2 + q0 q3)``).
"""
import numpy as np


def random_uniform_quaternion(rng=None) -> np.ndarray:
    """One Shoemake quaternion (h1, h2, h3, h4) from u1, u2, u3 ~ U[0, 1)."""
    u1, u2, u3 = (np.random.random(3) if rng is None else rng.random(3))
    s1, s2 = np.sqrt(1 - u1), np.sqrt(u1)
    t2, t3 = 2 * np.pi * u2, 2 * np.pi * u3
    return np.hstack([s1 * np.sin(t2), s1 * np.cos(t2), s2 * np.sin(t3), s2 * np.cos(t3)])


def random_uniform_quaternions(n: int, rng=None) -> np.ndarray:
    """(n, 4) quaternions, drawn one after another as the reference does.

    The n draws of three uniforms are taken in one call: numpy fills an
    array of uniforms from the same stream, in the same order, as n calls of
    three, so the quaternions equal the one-at-a-time ones bit for bit."""
    u = np.random.random((n, 3)) if rng is None else rng.random((n, 3))
    s1, s2 = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0])
    t2, t3 = 2 * np.pi * u[:, 1], 2 * np.pi * u[:, 2]
    return np.stack([s1 * np.sin(t2), s1 * np.cos(t2), s2 * np.sin(t3), s2 * np.cos(t3)],
                    axis=1).reshape(n, 4)


def get_quaternion(n, theta: float) -> np.ndarray:
    """Axis-angle -> unit quaternion (cos(theta/2), sin(theta/2) n)."""
    if not isinstance(n, np.ndarray):
        if isinstance(n, (list, tuple)):
            n = np.array(n)
        else:
            raise ValueError("rotation axis n must be an array-like object.")
    if n.shape != (3,):
        raise ValueError("rotation axis n must be a 3D vector.")
    n = n / np.linalg.norm(n)
    q = np.zeros(4, dtype=np.float64)
    q[0] = np.cos(theta / 2)
    q[1:] = np.sin(theta / 2) * n
    return q


def get_rotation_mattrices(quat: np.ndarray) -> np.ndarray:
    """(batch, 4) or (4,) quaternions -> (batch, 3, 3) float64 matrices."""
    quat = np.asarray(quat)
    if quat.ndim == 1:
        quat = quat[None, :]
    elif quat.ndim > 2:
        raise ValueError("ndim of quat must be 1 or 2")
    if quat.shape[1] != 4:
        raise ValueError("quaternion must be 4D vector")

    q0, q1, q2, q3 = (quat[:, i] for i in range(4))
    rot = np.empty((quat.shape[0], 3, 3), dtype=np.float64)
    rot[:, 0, 0] = 1. - 2. * (q2 * q2 + q3 * q3)
    rot[:, 0, 1] = 2. * (q1 * q2 + q0 * q3)
    rot[:, 0, 2] = 2. * (q1 * q3 - q0 * q2)
    rot[:, 1, 0] = 2. * (q1 * q2 - q0 * q3)
    rot[:, 1, 1] = 1. - 2. * (q1 * q1 + q3 * q3)
    rot[:, 1, 2] = 2. * (q0 * q1 + q2 * q3)
    rot[:, 2, 0] = 2. * (q0 * q2 + q1 * q3)
    rot[:, 2, 1] = 2. * (q2 * q3 - q0 * q1)
    rot[:, 2, 2] = 1. - 2. * (q1 * q1 + q2 * q2)
    return rot

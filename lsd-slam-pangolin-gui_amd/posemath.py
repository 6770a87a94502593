"""Pose algebra in numpy (FP64) for reporting: quaternion poses as
(qx, qy, qz, qw, tx, ty, tz[, scale]), SE3/SO3 log, and the twist-norm error
|log(A^-1 B)| used for every pose comparison."""
import numpy as np


def quat_to_R(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def to_matrix(p):
    p = np.asarray(p, np.float64)
    s = p[7] if len(p) > 7 else 1.0
    T = np.eye(4)
    T[:3, :3] = s * quat_to_R(p[:4])
    T[:3, 3] = p[4:7]
    return T


def so3_log(R):
    c = np.clip((np.trace(R) - 1.0) / 2.0, -1.0, 1.0)
    th = np.arccos(c)
    w = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    if th < 1e-7:
        return 0.5 * w
    if np.pi - th < 1e-6:
        # near pi: axis from the symmetric part
        A = (R + np.eye(3)) / 2.0
        k = int(np.argmax(np.diag(A)))
        ax = A[:, k] / np.sqrt(A[k, k])
        return th * ax
    return th / (2.0 * np.sin(th)) * w


def hat(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def se3_log(T):
    R = T[:3, :3] / np.cbrt(np.linalg.det(T[:3, :3]))
    w = so3_log(R)
    th = np.linalg.norm(w)
    W = hat(w)
    if th < 1e-7:
        Vinv = np.eye(3) - 0.5 * W + W @ W / 12.0
    else:
        Vinv = np.eye(3) - 0.5 * W + (1.0 / th ** 2) * (1 - (th * np.sin(th)) / (2 * (1 - np.cos(th)))) * (W @ W)
    return np.concatenate([Vinv @ T[:3, 3], w])


def twist_error(a, b):
    """|log(A^-1 B)| for SE3 poses (scale ignored)."""
    A, B = to_matrix(list(a)[:7]), to_matrix(list(b)[:7])
    return float(np.linalg.norm(se3_log(np.linalg.inv(A) @ B)))


def sim3_error(a, b):
    """(twist norm of the rigid part, |log(scale_a / scale_b)|) of A^-1 B."""
    A, B = to_matrix(a), to_matrix(b)
    D = np.linalg.inv(A) @ B
    s = np.cbrt(np.linalg.det(D[:3, :3]))
    Dr = D.copy()
    Dr[:3, :3] /= s
    return float(np.linalg.norm(se3_log(Dr))), float(abs(np.log(s)))

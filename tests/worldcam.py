"""Camera matrices for the world-streaming tests: cgmath's look_at_rh and perspective (textbook
forms) and camera.rs's OPENGL_TO_WGPU_MATRIX (camera.rs:9-14, given to cgmath's column-major
Matrix4::new), composed as Camera::chunk_generation_frustum_matrix (camera.rs:58-62).
Matrices are float32 arrays indexed [row, col]."""
import numpy as np

Z_NEAR = 2.0
Z_FAR = 16.0 * (4 + 1)
OPENGL_TO_WGPU = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0.5, 0.5, 0, 0, 0, 1], np.float64).reshape(4, 4).T


def look_at_rh(eye, target, up):
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    return np.array([[s[0], s[1], s[2], -s @ eye],
                     [u[0], u[1], u[2], -u @ eye],
                     [-f[0], -f[1], -f[2], f @ eye],
                     [0, 0, 0, 1]])


def perspective(fovy_deg, aspect, near, far):
    f = 1.0 / np.tan(np.radians(fovy_deg) / 2.0)
    return np.array([[f / aspect, 0, 0, 0],
                     [0, f, 0, 0],
                     [0, 0, (far + near) / (near - far), 2 * far * near / (near - far)],
                     [0, 0, -1, 0]])


def frustum(sub_pos, bearing, fovy_deg, aspect=16 / 9):
    """A chase camera behind the sub looking along its bearing."""
    b = np.asarray(bearing, np.float64)
    b /= np.linalg.norm(b)
    eye = np.asarray(sub_pos, np.float64) - 12.0 * b + np.array([0, 0, 3.0])
    view = look_at_rh(eye, np.asarray(sub_pos, np.float64), (0, 0, 1))
    return (OPENGL_TO_WGPU @ (perspective(fovy_deg, aspect, Z_NEAR, Z_FAR) @ view)).astype(np.float32)


def column_major(m):
    return np.ascontiguousarray(np.asarray(m, np.float32).T).ravel()

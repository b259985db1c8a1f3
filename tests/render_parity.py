"""Same-key radiance comparison against the FP64 oracle (test helper).

The GPU and the oracle draw identical random numbers (counter RNG keyed by
seed, pixel, sample, bounce), so a pixel can differ only where FP32 and FP64
geometry take different discrete decisions.  The oracle counts, per pixel,
its fragile decisions (crt_oracle_render_ex): closest hits within eps of a
triangle edge or of a closer triangle (after the 1/|cos| grazing stretch),
shadow rays whose occlusion flips when triangles grow or shrink by eps
(terminators), texel lookups at a texel boundary.  Every differing pixel
must carry one of these; eps is 16 FP32 ulp of the scene's diameter.
"""
import numpy as np


def fp32_eps(scene_info):
    """Decision band: 2^-20 x the scene's diameter (16 ulp of its FP32 coordinates)."""
    b = np.asarray(scene_info["bounds"], dtype=np.float64)
    return 2.0 ** -20 * float(np.max(b[3:] - b[:3]))


def classify(gpu_rgb, cpu_rgb, frag, rel=1e-3):
    """Per-pixel agreement within rel (plus 1e-5 absolute) and the classes of
    the pixels that differ."""
    diff = np.abs(gpu_rgb - cpu_rgb)
    close = np.all(diff <= rel * np.maximum(np.abs(cpu_rgb), 1e-3) + 1e-5, axis=-1)
    bad = ~close
    fc, fs, ft = frag[..., 0] > 0, frag[..., 1] > 0, frag[..., 2] > 0
    return {
        "pixels": int(close.size),
        "agree": float(close.mean()),
        "differ": int(bad.sum()),
        "closest": int((bad & fc).sum()),
        "shadow": int((bad & fs).sum()),
        "texel": int((bad & ft).sum()),
        "unexplained": int((bad & ~fc & ~fs & ~ft).sum()),
        "fragile_pixels": float((fc | fs | ft).mean()),
        "mean_rel": float(abs(gpu_rgb.mean() - cpu_rgb.mean()) / max(abs(cpu_rgb.mean()), 1e-30)),
        # systematic bias: the image mean over the pixels that agree (the
        # classified divergent paths excluded)
        "mean_rel_agreeing": float(abs(gpu_rgb[close].mean() - cpu_rgb[close].mean()) /
                                   max(abs(cpu_rgb[close].mean()), 1e-30)) if close.any() else 0.0,
    }

"""Writes the R0 fixtures of tests/golden/ from the reference's scene literals.

Run in the build container (the only place the reference sources exist):
    python tests/golden/make_r0_fixture.py [path/to/vulkan-raytraced-triangle/main.cpp]

r0_scene.json   the numbers the reference's scene is made of, read from main.cpp's text (vertex and index arrays,
                geometry and instance transforms, instance fields, camera block, traceRayEXT tmin / tmax, miss and
                hit-group colours, closest-hit condition, film size), each with the main.cpp line it was read from.
r0_film.npz     the expected RGBA32F film at every 8th pixel in x and y, computed in closed form from those
                numbers (every hit is at t = 10: origin z = 10, direction z = -1, quads in the z = 0 plane), and
                a mask of pixels within 1e-3 world units of a quad edge or of the barycentric-coloured triangle's
                diagonal (ties, excluded from comparisons).

Nothing is executed from the reference; its file is read as text only.
"""
import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SRC = "/root/reference/vulkan-raytraced-triangle/main.cpp"
STRIDE = 8


def _floats(text):
    return [float(x) for x in re.findall(r"-?\d+\.\d+|-?\d+", text.replace("f", ""))]


def _block(lines, start_pat, end_pat):
    """(first line number, text) from the line matching start_pat up to the next line matching end_pat."""
    for i, l in enumerate(lines):
        if re.search(start_pat, l):
            for j in range(i, len(lines)):
                if re.search(end_pat, lines[j]):
                    return i + 1, "\n".join(lines[i:j + 1])
    raise ValueError(f"pattern not found: {start_pat}")


def parse(src):
    lines = open(src).read().split("\n")
    out = {"source": "vulkan-raytraced-triangle/main.cpp"}

    ln, t = _block(lines, r"float vertices\[\]\[3\]", r"^\s*\};")
    out["vertices"] = {"line": ln, "value": np.array(_floats(t.split("=", 1)[1]), float).reshape(-1, 3).tolist()}
    ln, t = _block(lines, r"uint32_t indices\[\]", r";")
    out["indices"] = {"line": ln, "value": [int(x) for x in _floats(t.split("=", 1)[1])]}
    ln, t = _block(lines, r"VkTransformMatrixKHR geoTransforms\[\]", r"^\s*\};")
    out["geometry_transforms"] = {"line": ln,
                                  "value": np.array(_floats(t.split("=", 1)[1]), float).reshape(-1, 3, 4).tolist()}
    ln, t = _block(lines, r"VkTransformMatrixKHR insTransforms\[\]", r"^\s*\};")
    out["instance_transforms"] = {"line": ln,
                                  "value": np.array(_floats(t.split("=", 1)[1]), float).reshape(-1, 3, 4).tolist()}
    ln, t = _block(lines, r"VkAccelerationStructureInstanceKHR instance0", r"^\s*\};")
    custom = int(re.search(r"instanceCustomIndex\s*=\s*(\d+)", t).group(1))
    mask = int(re.search(r"\.mask\s*=\s*(0x[0-9A-Fa-f]+|\d+)", t).group(1), 0)
    sbt0 = int(re.search(r"instanceShaderBindingTableRecordOffset\s*=\s*(\d+)", t).group(1))
    flag = re.search(r"\.flags\s*=\s*(\w+)", t).group(1)
    ln1, t1 = _block(lines, r"instanceData\[1\]\.instanceShaderBindingTableRecordOffset", r";")
    sbt1 = int(re.search(r"=\s*(\d+)", t1).group(1))
    out["instances"] = {"line": ln, "custom_index": [custom, custom], "mask": [mask, mask],
                        "sbt_record_offset": [sbt0, sbt1], "sbt_offset_line": ln1,
                        "flags": [flag, flag]}
    ln, t = _block(lines, r"\*\(Data\*\) dst = \{", r";")
    cam = _floats(t.split("=", 1)[1])
    out["camera"] = {"line": ln, "position": cam[:3], "y_fov_degree": cam[3]}
    ln, t = _block(lines, r"g\.cameraPos, [\d.]+, rayDir, [\d.]+", r"\)")
    m = re.search(r"g\.cameraPos,\s*([\d.]+),\s*rayDir,\s*([\d.]+)", t)
    out["trace"] = {"line": ln, "tmin": float(m.group(1)), "tmax": float(m.group(2))}
    ln, t = _block(lines, r"hitValue = vec3\(0\.0, 0\.0, 0\.2\)", r";")
    out["miss_color"] = {"line": ln, "value": _floats(t.split("vec3", 1)[1])}
    cols = []
    first = None
    for i, l in enumerate(lines):
        m = re.search(r"\(HitgCustomData\*\s*\)\(dst \+ hitgOffset \+ (\d) \* hitgStride \+ handleSize\) = \{([^}]*)\}", l)
        if m:
            first = first or i + 1
            cols.append((int(m.group(1)), _floats(m.group(2))))
    out["hit_colors"] = {"line": first, "value": [c for _, c in sorted(cols)]}
    ln, t = _block(lines, r"if \(gl_PrimitiveID ==", r"\{")
    cond = {k: int(v) for k, v in re.findall(r"(gl_\w+)\s*==\s*(\d+)", t)}
    out["closest_hit_match"] = {"line": ln, "value": cond}
    w = re.search(r"const uint32_t WIDTH = (\d+);", "\n".join(lines))
    h = re.search(r"const uint32_t HEIGHT = (\d+);", "\n".join(lines))
    out["film"] = {"width": int(w.group(1)), "height": int(h.group(1))}
    return out


def closed_form_film(sc, stride=STRIDE):
    """Expected film at pixels (stride*i, stride*j) and the tie mask, from the scene numbers alone."""
    W, H = sc["film"]["width"], sc["film"]["height"]
    cam = sc["camera"]
    ay = math.tan(math.radians(cam["y_fov_degree"]) / 2)
    ax = ay * W / H
    x = np.arange(0, W, stride)
    y = np.arange(0, H, stride)
    nx = (x + 0.5) / W * 2 - 1
    ny = (y + 0.5) / H * 2 - 1
    z0 = cam["position"][2]  # the quads lie in z = 0; direction z = -1, so every hit is at t = z0
    X = np.broadcast_to(cam["position"][0] + z0 * ax * nx[None, :], (len(y), len(x)))
    Y = np.broadcast_to(cam["position"][1] - z0 * ay * ny[:, None], (len(y), len(x)))
    img = np.zeros((len(y), len(x), 4))
    img[..., :3] = sc["miss_color"]["value"]
    tie = np.zeros((len(y), len(x)), bool)
    V = np.array(sc["vertices"]["value"])
    idx = np.array(sc["indices"]["value"]).reshape(-1, 3)
    match = sc["closest_hit_match"]["value"]
    for inst, T in enumerate(sc["instance_transforms"]["value"]):
        for g, G in enumerate(sc["geometry_transforms"]["value"]):
            off = np.array(T)[:, 3] + np.array(G)[:, 3]  # translations only in this scene
            P = V + off
            rec = sc["instances"]["sbt_record_offset"][inst] + g  # sbt offset + geometry index x stride 1
            for prim, (i0, i1, i2) in enumerate(idx):
                a, b, c = P[i0, :2], P[i1, :2], P[i2, :2]
                den = (b[1] - c[1]) * (a[0] - c[0]) + (c[0] - b[0]) * (a[1] - c[1])
                l0 = ((b[1] - c[1]) * (X - c[0]) + (c[0] - b[0]) * (Y - c[1])) / den
                l1 = ((c[1] - a[1]) * (X - c[0]) + (a[0] - c[0]) * (Y - c[1])) / den
                l2 = 1 - l0 - l1
                inside = (l0 >= 0) & (l1 >= 0) & (l2 >= 0)
                # distance to the triangle's edges, for the tie mask
                for p, q in ((a, b), (b, c), (c, a)):
                    e = q - p
                    d = np.abs(e[0] * (Y - p[1]) - e[1] * (X - p[0])) / np.hypot(*e)
                    tie |= (d < 1e-3) & (np.abs(l0) < 1.001) & (np.abs(l1) < 1.001) & (np.abs(l2) < 1.001)
                is_match = (prim == match["gl_PrimitiveID"] and inst == match["gl_InstanceID"]
                            and sc["instances"]["custom_index"][inst] == match["gl_InstanceCustomIndexEXT"]
                            and g == match["gl_GeometryIndexEXT"])
                if is_match:
                    # hitAttributeEXT (u, v): P = (1-u-v) v0 + u v1 + v v2 -> colour (1-u-v, u, v)
                    col = np.stack([l0, l1, l2], -1)
                    img[inside, :3] = col[inside]
                else:
                    img[inside, :3] = sc["hit_colors"]["value"][rec]
    return img.astype(np.float32), tie


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    sc = parse(src)
    with open(os.path.join(HERE, "r0_scene.json"), "w") as f:
        json.dump(sc, f, indent=1)
        f.write("\n")
    img, tie = closed_form_film(sc)
    np.savez_compressed(os.path.join(HERE, "r0_film.npz"), stride=np.int32(STRIDE), rgba=img, tie=tie)
    print(f"wrote r0_scene.json and r0_film.npz ({img.shape[1]}x{img.shape[0]} samples, {int(tie.sum())} ties)")


if __name__ == "__main__":
    main()

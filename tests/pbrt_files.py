"""pbrt-v2 scene files written from the test scenes (for the import tests)."""
import numpy as np


def _f(x):
    """Shortest text that reads back (std::stof / std::stod) as the same float32."""
    return repr(float(np.float32(x)))


def _list(a):
    return "[" + " ".join(_f(x) for x in np.asarray(a).reshape(-1)) + "]"


def _ints(a):
    return "[" + " ".join(str(int(x)) for x in np.asarray(a).reshape(-1)) + "]"


def mesh_shape(m):
    s = 'Shape "trianglemesh" "integer indices" %s "point P" %s' % (_ints(m["tris"]), _list(m["verts"]))
    if m.get("normals") is not None:
        s += ' "normal N" %s' % _list(m["normals"])
    return s + "\n"


def room_text(s, width, height, fov_deg):
    """A direct-lighting room (workloads.scene_room layout: one mesh per primitive, lights last) as a
    pbrt file: LookAt camera from s["camera"], one attribute block per primitive with its matte /
    plastic material, area lights with their radiance."""
    origin, d, up, _ = s["camera"]
    look = np.asarray(origin, np.float64) + np.asarray(d, np.float64)
    out = ["# test room\n",
           "LookAt %s %s %s\n" % (" ".join(_f(x) for x in origin), " ".join(_f(x) for x in look),
                                  " ".join(_f(x) for x in up)),
           'Camera "perspective" "float fov" [%s]\n' % _f(fov_deg),
           'Film "image" "integer xresolution" [%d] "integer yresolution" [%d]\n' % (width, height),
           "WorldBegin\n"]
    lights = {int(p): i for i, p in enumerate(s["light_prim"])}
    for p, m in enumerate(s["meshes"]):
        out.append("AttributeBegin\n")
        if s["mat"][p] == 1:
            out.append('Material "plastic" "rgb Kd" %s "rgb Ks" %s "float roughness" %s\n' %
                       (_list(s["kd"][p]), _list(s["ks"][p]), _f(s["rough"][p])))
        else:
            out.append('Material "matte" "rgb Kd" %s\n' % _list(s["kd"][p]))
        if p in lights:
            out.append('AreaLightSource "diffuse" "rgb L" %s\n' % _list(s["light_radiance"][lights[p]]))
        if m.get("invert_normals"):
            out.append("ReverseOrientation\n")
        out.append(mesh_shape(m))
        out.append("AttributeEnd\n")
    out.append("WorldEnd\n")
    return "".join(out)

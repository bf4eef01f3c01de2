"""Minimal reader of the sba text format used by the reference's data files
(points: `X Y Z nframes  frame x y ...`; cameras: `fu u0 v0 ar s  q0 q1 q2 q3
tx ty tz`), for building test inputs.  Comment lines start with '#'."""
import numpy as np


def read_points(path):
    pts, tracks = [], []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            v = s.split()
            X = [float(x) for x in v[:3]]
            nf = int(v[3])
            fr = []
            for k in range(nf):
                j = int(v[4 + 3 * k])
                fr.append((j, float(v[5 + 3 * k]), float(v[6 + 3 * k])))
            pts.append(X)
            tracks.append(fr)
    return np.array(pts), tracks


def read_cams(path):
    rows = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            rows.append([float(x) for x in s.split()])
    return np.array(rows)


def vmask_of(tracks, n_cams):
    vm = np.zeros((len(tracks), n_cams), np.int8)
    for i, fr in enumerate(tracks):
        for j, _, _ in fr:
            vm[i, j] = 1
    return vm

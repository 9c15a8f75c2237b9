"""Interaction dumps in the reference's text format (src/interaction/include/exaDEM/itools/writer.hpp:35-115).

One row per interaction in contact (non-zero force or dn < 0), owner side, 21 comma-separated columns:
owner id, partner id (wall index for walls), owner sub-feature, partner sub-feature, type, dn, contact
point (3), normal force f - ft (3), tangential force ft (3), owner position (3), partner position (3;
zeros for walls). Numbers are written as a default C++ ostream writes doubles (6 significant digits,
%g). The file is <directory>/ExaDEMAnalyses/<basename><timestep>/<basename><timestep>_<rank>.txt
(write_file, writer.hpp:95-115), as contact_sphere writes it every analysis_interaction_dump_frequency
steps (contact_force_sphere.cu.in:190-199).
"""
import os

import numpy as np

NCOLS = 21


def _num(x):
    return "%g" % x


def format_rows(id_i, id_j, sub_i, sub_j, itype, values):
    """Lines of the dump; values is [rows][16] (dn, contact point, f - ft, ft, owner position, partner
    position)."""
    out = []
    for k in range(len(id_i)):
        v = values[k]
        cols = [str(int(id_i[k])), str(int(id_j[k])), str(int(sub_i[k])), str(int(sub_j[k])), str(int(itype[k]))]
        cols += [_num(float(x)) for x in v]
        out.append(",".join(cols))
    return out


def dump_path(directory, basename, timestep, rank=0):
    name = f"{basename}{timestep}"
    return os.path.join(directory, "ExaDEMAnalyses", name, f"{name}_{rank}.txt")


def write_dump(ctx, directory, basename, timestep, rank=0):
    """Writes the contact state recorded by the context's last force pass (spheres or polyhedra; the
    context must record: ctx.set_contact_record(True) before the pass). Returns the file path and the
    row count."""
    ii, jj, si, sj, tt, vals = ctx.contact_state(subs=True)
    lines = format_rows(ii, jj, si, sj, tt, vals)
    path = dump_path(directory, basename, timestep, rank)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        for line in lines:
            f.write(line + "\n")
    return path, len(lines)


def read_dump(path):
    """Parses a dump file: (id_i, id_j, sub_i, sub_j, type) as int64 [rows][5] and the 16 values as
    float64 [rows][16]."""
    keys, vals = [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            cols = line.split(",")
            if len(cols) != NCOLS:
                raise ValueError(f"{path}: {len(cols)} columns, expected {NCOLS}")
            keys.append([int(c) for c in cols[:5]])
            vals.append([float(c) for c in cols[5:]])
    return np.array(keys, np.int64).reshape(-1, 5), np.array(vals, np.float64).reshape(-1, 16)

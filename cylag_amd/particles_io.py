"""Particle result files: the TXT and VTK writers the Solver calls every
output_printout_period steps, the stranded-particle file, and ParticlesIO,
which reads TXT results back.

TXT layout (cylag/io/write_particles.pyx:18-50, cylag/io/writer_txt.pyx:25-180):
  cylag particles result file
  Variables: Tags, Xp, Yp, Zp, Up, Vp, Wp, [Us, Vs, Ws, ] [dp, ] [depth, ]
then per record a header line "Time = %.8f, Iter = %16d, Npart = %16d" and one
line per active particle: tag, positions (%.16e; in 2D the z column is 0),
velocities (%.8e; in 2D the w column is 0), fluid velocity seen for LSM2,
diameter for non-monodisperse sets, depth when computed.
"""
import os

import numpy as np


def prepare_output(pset, outputdir, output_file_format, output_file_name, model=1):
    """Create the output directory and write the TXT header (overwrites)."""
    os.makedirs(outputdir, exist_ok=True)
    if output_file_format in ('txt', 'all'):
        with open(os.path.join(outputdir, output_file_name + '.txt'), 'w') as f:
            f.write('cylag particles result file\n')
            f.write('Variables: Tags, Xp, Yp, Zp, Up, Vp, Wp, ')
            if model >= 2:
                f.write('Us, Vs, Ws, ')
            if pset.additional_velocity:
                f.write('Ua, Va, Wa, ')
            if pset.additional_force:
                f.write('Fax, Fay, Faz, ')
            if pset.diameter_distribution != 'monodisperse':
                f.write('dp, ')
            if pset.compute_depth:
                f.write('depth, ')
            f.write('\n')


def write_particles(pset, time, iteration, params):
    """One output record of the active particles (write_particles.pyx:52-91):
    a VTK file per record ('vtk', 'all') and a record appended to the TXT file
    ('txt', 'all')."""
    fmt = params.output_file_format
    if fmt not in ('txt', 'vtk', 'all'):
        raise ValueError("output_file_format {!r}: the writers are 'txt', 'vtk' and 'all'".format(fmt))
    if fmt in ('vtk', 'all'):
        write_vtk(pset, time, iteration, params.output_rep, params.output_file_name, params.model)
    if fmt in ('txt', 'all'):
        write_txt(pset, time, iteration, params)


def write_vtk(pset, time, iteration, outputdir, file_name, model):
    """Legacy-VTK unstructured grid of vertex cells, <file_name>_<iteration>.vtk
    (writer_vtk.pyx:7-304): time and cycle field data, points (z = 0 in 2D),
    then per point the pool index, tag, elevation (3D), velocity, diameter
    (non-monodisperse), fluid velocity seen (model >= 2) and depth (3D, when
    computed).  Values are written with Python's shortest round-trip repr."""
    dim = pset.dim
    act = np.nonzero(np.asarray(pset.inactive) == 0)[0]
    npa = act.size
    pos, vel = np.asarray(pset.position), np.asarray(pset.velocity)
    tag = np.asarray(pset.tag)

    def vec(a):
        return "".join("{} {} {}\n".format(a[j, 0], a[j, 1], a[j, 2] if dim == 3 else 0.) for j in act)

    def sca(a):
        return "".join("{}\n".format(a[j]) for j in act)

    out = ["# vtk DataFile Version 2.0\n", "particle positions (Points {}D)\n".format(dim), "ASCII\n",
           "DATASET UNSTRUCTURED_GRID\n", "FIELD FIELDDATA 2\n", "TIME 1 1 DOUBLE\n", "{}\n".format(time),
           "CYCLE 1 1 INT\n", "{}\n".format(iteration), "POINTS {} FLOAT\n".format(npa), vec(pos),
           "CELLS {} {}\n".format(npa, 2 * npa), "".join("1 {}\n".format(c) for c in range(npa)),
           "CELL_TYPES {}\n".format(npa), "1\n" * npa, "POINT_DATA {}\n".format(npa),
           "SCALARS Index INT\n", "LOOKUP_TABLE default\n", "".join("{}\n".format(j) for j in act),
           "SCALARS tag INT\n", "LOOKUP_TABLE default\n", "".join("{}\n".format(int(tag[j])) for j in act)]
    if dim == 3:
        out += ["SCALARS elevation FLOAT\n", "LOOKUP_TABLE default\n", sca(pos[:, 2])]
    out += ["VECTORS velocity FLOAT\n", vec(vel)]
    if pset.diameter_distribution != 'monodisperse':
        out += ["SCALARS diameter FLOAT\n", "LOOKUP_TABLE default\n", sca(np.asarray(pset.diameter))]
    if model >= 2:
        out += ["VECTORS fluid_velocity_seen FLOAT\n", vec(np.asarray(pset.fluid_velocity_seen))]
    if dim == 3 and pset.compute_depth:
        out += ["SCALARS depth FLOAT\n", "LOOKUP_TABLE default\n", sca(np.asarray(pset.depth))]
    with open(os.path.join(outputdir, file_name + "_{}.vtk".format(iteration)), "w") as f:
        f.write("".join(out))


def write_txt(pset, time, iteration, params):
    """Append one TXT record (writer_txt.pyx:25-180)."""
    dim = pset.dim
    inactive = np.asarray(pset.inactive)
    act = np.nonzero(inactive == 0)[0]
    pos, vel = np.asarray(pset.position), np.asarray(pset.velocity)
    us = np.asarray(pset.fluid_velocity_seen) if params.model >= 2 else None
    diam = np.asarray(pset.diameter) if pset.diameter_distribution != 'monodisperse' else None
    depth = np.asarray(pset.depth) if (dim == 3 and pset.compute_depth) else None
    tag = np.asarray(pset.tag)
    out = []
    out.append("Time = {:.8f}, ".format(time))
    out.append("Iter = {:16d}, ".format(iteration))
    out.append("Npart = {:16d}\n".format(pset.npart_active))
    for j in act:
        row = ['{:16d}, '.format(int(tag[j])), '{:.16e}, '.format(pos[j, 0]), '{:.16e}, '.format(pos[j, 1])]
        row.append('{:.16e}, '.format(pos[j, 2]) if dim == 3 else '{:1.0f}, '.format(0.))
        row += ['{:.8e}, '.format(vel[j, 0]), '{:.8e}, '.format(vel[j, 1])]
        row.append('{:.8e}, '.format(vel[j, 2]) if dim == 3 else '{:1.0f}, '.format(0.))
        if us is not None:
            row += ['{:.8e}, '.format(us[j, 0]), '{:.8e}, '.format(us[j, 1])]
            row.append('{:.8e}, '.format(us[j, 2]) if dim == 3 else '{:1.0f}, '.format(0.))
        if diam is not None:
            row.append('{:.8e}, '.format(diam[j]))
        if depth is not None:
            row.append('{:.8e}, '.format(depth[j]))
        row.append('\n')
        out.append(''.join(row))
    with open(os.path.join(params.output_rep, params.output_file_name + '.txt'), 'a') as f:
        f.write(''.join(out))


def write_stranded_particles_txt(outputdir, file_name, stranded_positions, stranded_count):
    """Stranded positions recorded since the last flush (write_particles.pyx:339-388):
    the file is rewritten at each flush that has records."""
    if stranded_count <= 0:
        return
    os.makedirs(outputdir, exist_ok=True)
    rows = np.asarray(stranded_positions)[:stranded_count]
    with open(os.path.join(outputdir, file_name + '.txt'), 'w') as f:
        f.write('cylag stranded particles file\nVariables: Xp, Yp, Zp\n')
        f.write(''.join('{:.16e}, {:.16e}, {:.16e}\n'.format(r[0], r[1], r[2]) for r in rows))


def write_deposited_particles_txt(outputdir, file_name, positions, iterations, count):
    """Deposits on the bed recorded since the last flush (bottom_deposition, an
    extension): the stranded file's layout with the iteration of each deposit;
    the file is rewritten at each flush that has records."""
    if count <= 0:
        return
    os.makedirs(outputdir, exist_ok=True)
    rows = np.asarray(positions)[:count]
    its = np.asarray(iterations)[:count]
    with open(os.path.join(outputdir, file_name + '.txt'), 'w') as f:
        f.write('cylag deposited particles file\nVariables: Xp, Yp, Zp, Iteration\n')
        f.write(''.join('{:.16e}, {:.16e}, {:.16e}, {:d}\n'.format(r[0], r[1], r[2], int(i))
                        for r, i in zip(rows, its)))


class ParticlesIO:
    """Records of a particle result file (particles_io.pyx:26-120, 453-648):
    per record the time, iteration and arrays of tags, positions and velocities
    (and the optional columns the header lists)."""

    def __init__(self, times, iterations, tags, xp, yp, zp, up=None, vp=None, wp=None, us=None, vs=None, ws=None,
                 diameter=None, depth=None):
        self.times = np.asarray(times, dtype='d')
        self.iterations = np.asarray(iterations, dtype=np.int64)
        self.ntimes = self.times.shape[0]
        self.tags = tags
        self.xp, self.yp, self.zp = xp, yp, zp
        self.up, self.vp, self.wp = up, vp, wp
        self.us, self.vs, self.ws = us, vs, ws
        self.diameter = diameter
        self.depth = depth
        self.npart = np.array([len(t) for t in tags], dtype=np.int64)

    @staticmethod
    def from_cylag_txt(file_name):
        with open(file_name, 'r') as f:
            f.readline()
            header = f.readline()
            cols = [c.strip() for c in header.split(':', 1)[1].split(',') if c.strip()]
            recs = []
            line = f.readline()
            while line:
                if line.startswith('Time = '):
                    parts = line.replace('=', ',').split(',')
                    t, it, n = float(parts[1]), int(parts[3]), int(parts[5])
                    rows = [f.readline() for _ in range(n)]
                    data = np.array([[float(v) for v in r.split(',')[:-1]] for r in rows], dtype='d') \
                        if n else np.zeros((0, len(cols)))
                    recs.append((t, it, data))
                line = f.readline()
        idx = {c: i for i, c in enumerate(cols)}
        col = lambda name: [d[:, idx[name]] for _, _, d in recs] if name in idx else None
        tags = [d[:, 0].astype(np.int32) for _, _, d in recs]
        return ParticlesIO([r[0] for r in recs], [r[1] for r in recs], tags, col('Xp'), col('Yp'), col('Zp'),
                           col('Up'), col('Vp'), col('Wp'), col('Us'), col('Vs'), col('Ws'), col('dp'),
                           col('depth'))

    def get_trajectory(self, tag):
        """(k, 3) positions of one tagged particle over the records where it exists."""
        pts = []
        for k in range(self.ntimes):
            w = np.nonzero(self.tags[k] == tag)[0]
            if w.size:
                i = w[0]
                pts.append((self.xp[k][i], self.yp[k][i], self.zp[k][i]))
        return np.array(pts, dtype='d').reshape(-1, 3)

    def get_trajectories(self, tags):
        return [self.get_trajectory(t) for t in tags]

"""Boundary and control-section statistics files
(cylag/io/write_statistics.pyx:24-107): one file per open boundary or section,
``<name>_<k>.txt`` (k from 1), a two-line header, then one line per step,
"time, count" as "{:.8f}, {:16d}  "."""
import os


def prepare_secstats_output(bnd_statistics=False, nopen=0, outputdir="particles", bnd_file_name="bnd_stats",
                            cs_statistics=False, ncsec=0, cs_file_name="cs_stats"):
    """Create the directory and write the headers (overwrites existing files)."""
    os.makedirs(outputdir, exist_ok=True)
    if bnd_statistics:
        for k in range(1, nopen + 1):
            with open(os.path.join(outputdir, "{}_{}.txt".format(bnd_file_name, k)), "w") as f:
                f.write("cylag boundary statistics file\nVariables: Time, Npart\n")
    if cs_statistics:
        for k in range(1, ncsec + 1):
            with open(os.path.join(outputdir, "{}_{}.txt".format(cs_file_name, k)), "w") as f:
                f.write("cylag control section file\nVariables: Time, Npart\n")


def write_secstats_buffer(times, data, nsteps, nsec, outputdir, file_name):
    """Append nsteps buffered rows (times[j], data[j, k - 1]) to each file k."""
    for k in range(1, nsec + 1):
        with open(os.path.join(outputdir, "{}_{}.txt".format(file_name, k)), "a") as f:
            f.write("".join("{:.8f}, {:16d}  \n".format(times[j], int(data[j, k - 1])) for j in range(nsteps)))

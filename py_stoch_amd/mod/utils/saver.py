"""Text output of simulation results (reference: mod/utils/saver.py:8-67).

One ``numpy.savetxt(fmt='%1u')`` file per recorded species under
``<cwd>/<sim_name>/x-<x0...>_c-<c...>/<sim_type>_species<k>_<y-m-d-h-min>``
with the run settings in the header.
"""
import os
from datetime import datetime

import numpy


def save_results(sim_result, arg_obj, sim_name, sim_type="TL", out_root=None):
    root = os.getcwd() if out_root is None else out_root
    now = datetime.now()
    dt = "{0}-{1}-{2}-{3}-{4}".format(now.year, now.month, now.day, now.hour, now.minute)

    x0 = "x"
    try:
        for i in range(len(arg_obj.x_0)):
            x0 += "-" + str(arg_obj.x_0[i])
    except TypeError:
        x0 += "-" + str(arg_obj.x_0)

    c = "_c"
    try:
        for i in range(len(arg_obj.c)):
            c += "-" + str(arg_obj.c[i]).rstrip(".0")
    except TypeError:
        c += "-" + str(arg_obj.c).rstrip(".0")

    out_dir = os.path.join(root, sim_name, x0 + c) + "/"
    try:
        os.makedirs(out_dir)
    except OSError:
        print("folder already exists")

    header = make_header(arg_obj)
    written = []
    if sim_result.ndim == 3:
        for k in range(len(sim_result)):
            filename = out_dir + sim_type + "_species" + str(k) + "_" + dt
            numpy.savetxt(filename, sim_result[k], fmt="%1u", header=header)
            written.append(filename)
    elif sim_result.ndim == 2:
        for k in range(len(sim_result[0])):
            filename = out_dir + sim_type + "_species" + str(k) + "_" + dt
            numpy.savetxt(filename, sim_result[:, k], fmt="%1u", header=header)
            written.append(filename)
    return written


def make_header(args):
    return """
output_species: {0}
duration: {1}
num_output_time_points: {2}
num_simulations: {3}
species_init: {4}
parameters:{5}
""".format(str(args.E), str(args.t_max), str(args.ita), str(args.U), str(args.x_0), str(args.c))

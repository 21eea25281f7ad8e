"""Helpers shared by the GPU tests."""
import numpy as np

import mdcore as md
import pyoracle as po
import systems


def engine_pass(sys, flags=md.engine_flag_gpu):
    """engine_init .. engine_start, then one nonbonded pass on the GPU."""
    e = systems.build_engine(sys, flags)
    e.shuffle()
    e.nonbond()
    return e


def rel_rmse(f, ref):
    return float(np.sqrt(np.sum((f - ref) ** 2) / np.sum(ref ** 2)))


def oracle_for(sys):
    return po.from_system(sys)

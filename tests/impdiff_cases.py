"""Case geometries for the implicit-diffusion tests (the reference's _IMPDIFF
build), written as dns.nml / blocks.nml like the other cases.

channel  : the periodic channel (x, z periodic, y tanh-stretched walls);
           the wall-normal velocity is staggered along the line axis
cavity   : the lid-driven cavity on a uniform grid, all walls (u = lid on
           z+): staggered Dirichlet pairs along transformed axes (DST-I over
           n-1 points) and along the FFT axis with FFT_AXIS = 3
duct     : inflow at y = 0 (Poiseuille profile across the x walls), outflow
           at y = Ly (velocity N, pressure D), walls in x, z periodic, y
           stretched: the inflow plane scales the boundary terms
box      : a periodic box along every axis (every axis transformed)
channel_scalar : the channel with a scalar between walls at 0 and 1,
           forced to a bulk value (is_sforced) or not
"""
import numpy as np

import cases

F1 = ("F", 1)


def channel(n=(32, 24, 16)):
    return [cases.channel_block(n)], dict(bforce=(12 / 5600, 0, 0))


def cavity(n=(16, 12, 20)):
    return [cases.cavity_block(n, lid=1.0)], {}


def duct(n=(16, 24, 8)):
    blk = cases._block(n, (0, 0, 0), (1.0, 3.0, 0.5),
                       vel=[["D", "D"], ["D", "N"], [F1, F1]], pre=[["N", "N"], ["N", "D"], [F1, F1]],
                       gr=(0.0, 1.5, 0.0), vals={(1, 1, 0): 1.0}, inflow={(1, 0): 1})
    return [blk], {}


def box(n=(16, 8, 12)):
    return [cases.triperiodic_block(n, lengths=(2 * np.pi, 2 * np.pi, 2 * np.pi))], {}


def channel_scalar(n=(32, 24, 16), forced=True):
    blocks, dns = channel(n)
    blocks[0]["cbcscal"] = [[["F", "F"], ["D", "D"], ["F", "F"]]]
    blocks[0]["bcscal"] = [[[1.0, 1.0], [0.0, 1.0], [1.0, 1.0]]]
    dns.update(nscal=1, gacc=(0.0, -1.0, 0.0),
               scalar=dict(alphai=[80.0], beta=0.3, iniscal=["zer"], ssource=[0.2], is_sforced=[forced],
                           scalf=[0.5]))
    return blocks, dns


GEOMS = {"channel": channel, "cavity": cavity, "duct": duct, "box": box, "channel_scalar": channel_scalar,
         "channel_scalar_free": lambda: channel_scalar(forced=False)}

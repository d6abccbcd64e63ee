"""Synthetic shallow-water workloads with recorded seeds.

* ``dam_break_channel`` -- the 1000 m x 1000 m rectangular_cross dam break with
  a sloping bed and Gaussian bump, Manning n = 0.03, reflective walls (the
  CPU-runnable FP64 workload; 100 x 100 cells = 40,000 triangles; 1000 x 1000
  cells = 4,000,000 triangles).
* ``delaunay_terrain`` -- seeded Delaunay triangulation of uniform points in a
  10 km x 10 km square (500k points give ~1.0M triangles) with 5-octave fBm
  terrain of amplitude 20 m, stage max(z, 5 m) plus a 3 m mound of radius
  500 m, Manning n = 0.025, transmissive boundaries; triangles stored in Morton
  order, random order or the triangulator's order.
* ``merimbula`` -- the reference's Merimbula scenario
  (tests/CUDA/merimbula/merimbula.py:44-60, 155-165): mesh
  ``merimbula_43200.tsh``, flat bed, stage 1 m plus 2 m for
  756000 < x < 756500, reflective walls on every tag.
"""
from __future__ import annotations

import os

import numpy as np

from .domain import Dirichlet_boundary, Domain, Reflective_boundary, Transmissive_boundary
from .mesh import Mesh, delaunay_jittered, delaunay_random, read_tsh, rectangular_cross

__all__ = ["dam_break_channel", "delaunay_terrain", "fbm", "merimbula", "MERIMBULA_TSH"]

MERIMBULA_TSH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "tests", "golden", "merimbula_43200.tsh.gz")


def dam_break_channel(m=100, n=100, friction=0.03, dam_x=500.0, high=2.0, low=0.5,
                      bump=0.5, sigma=100.0, precision="fp64", **params):
    points, tris, boundary = rectangular_cross(m, n, 1000.0, 1000.0)
    mesh = Mesh(points, tris, boundary=boundary)
    d = Domain(mesh=mesh, precision=precision, **params)

    def bed(x, y):
        r2 = (x - 500.0) ** 2 + (y - 500.0) ** 2
        return x / 1000.0 + bump * np.exp(-r2 / (2.0 * sigma * sigma))

    def stage(x, y):
        return np.maximum(np.where(x < dam_x, high, low), bed(x, y))

    d.set_quantity("elevation", bed)
    d.set_quantity("stage", stage)
    d.set_quantity("friction", friction, location="centroids")
    d.set_boundary({tag: Reflective_boundary(d) for tag in mesh.get_tag_names()})
    return d


def fbm(x, y, seed=2024, octaves=5, amplitude=20.0, base_wavelength=5000.0, modes=6):
    """Seeded fractional-Brownian-motion-like terrain: per octave ``modes``
    plane waves with random directions and phases, amplitude halving and
    frequency doubling per octave; normalised so the sum of octave amplitudes
    equals ``amplitude``."""
    rng = np.random.default_rng(seed)
    z = np.zeros_like(np.asarray(x, dtype=np.float64))
    total = 0.0
    for o in range(octaves):
        a = 0.5 ** o
        k = 2.0 * np.pi * (2.0 ** o) / base_wavelength
        theta = rng.uniform(0.0, 2.0 * np.pi, modes)
        phase = rng.uniform(0.0, 2.0 * np.pi, modes)
        for j in range(modes):
            z = z + (a / np.sqrt(modes)) * np.cos(k * (np.cos(theta[j]) * x + np.sin(theta[j]) * y) + phase[j])
        total += a
    return amplitude * z / total


def delaunay_terrain(npoints=500000, order="morton", seed=12345, terrain_seed=2024,
                     extent=10000.0, friction=0.025, base_stage=5.0, mound=3.0,
                     mound_radius=500.0, precision="fp64", build="host", **params):
    """``build`` ('host' or 'device'): where the mesh geometry is built (the
    same arrays either way)."""
    points, tris = delaunay_random(npoints, extent=(extent, extent), seed=seed)
    mesh = Mesh(points, tris, build=build)
    if order == "morton":
        mesh = mesh.morton_renumbered()
    elif order == "random":
        mesh = mesh.random_renumbered(seed=seed + 1)
    elif order not in (None, "none"):
        raise ValueError("order must be 'morton', 'random' or 'none'")
    d = Domain(mesh=mesh, precision=precision, **params)
    cx = cy = extent / 2.0

    def bed(x, y):
        return fbm(x, y, seed=terrain_seed, base_wavelength=extent / 2.0)

    def stage(x, y):
        r2 = (x - cx) ** 2 + (y - cy) ** 2
        return np.maximum(bed(x, y), base_stage) + np.where(r2 < mound_radius ** 2, mound, 0.0)

    d.set_quantity("elevation", bed)
    d.set_quantity("stage", stage)
    d.set_quantity("friction", friction, location="centroids")
    d.set_boundary({tag: Transmissive_boundary(d) for tag in mesh.get_tag_names()})
    return d


def merimbula(path=MERIMBULA_TSH, x0=756000.0, x1=756500.0, h=2.0, **params):
    points, tris, segments = read_tsh(path)
    mesh = Mesh(points, tris, segments=segments)
    d = Domain(mesh=mesh, **params)
    d.set_quantity("stage", lambda x, y: h * ((x > x0) & (x < x1)) + 1.0)
    d.set_boundary({tag: Reflective_boundary(d) for tag in mesh.get_tag_names()})
    return d


def tag_square_sides(mesh, lx, ly):
    """Tag the boundary edges of a mesh of the rectangle [0, lx] x [0, ly] as
    'left', 'right', 'bottom' or 'top' by the side their midpoint lies on."""
    t, p = mesh.triangles, mesh.points
    k, e = mesh.boundary_cells, mesh.boundary_edges
    a = p[t[k, (e + 1) % 3]]
    b = p[t[k, (e + 2) % 3]]
    mx, my = 0.5 * (a[:, 0] + b[:, 0]), 0.5 * (a[:, 1] + b[:, 1])
    dist = np.stack([mx, lx - mx, my, ly - my], axis=1)
    names = np.array(["left", "right", "bottom", "top"], dtype=object)
    tags = names[np.argmin(dist, axis=1)]
    mesh.boundary_tags = tags
    mesh.boundary = {(int(kk), int(ee)): tag for kk, ee, tag in zip(k, e, tags)}
    return mesh


def ensemble_member(triangles, seed, manning, dam_height, inflow, extent=1000.0, precision="fp64", **params):
    """One scenario of the batched ensemble: a seeded Delaunay mesh (jittered
    grid points) of about ``triangles`` triangles on a square of side
    ``extent``, bed sloping 1:1000
    with a Gaussian bump, a dam break of ``dam_height`` m against 0.5 m, Manning
    n ``manning``, an inflow boundary on the left (Dirichlet: stage dam_height,
    x-momentum ``inflow`` m^2/s), transmissive on the right, reflective walls
    top and bottom."""
    points, tris = delaunay_jittered(max(int(triangles) // 2, 50), extent=(extent, extent), seed=seed)
    mesh = tag_square_sides(Mesh(points, tris), extent, extent)
    d = Domain(mesh=mesh, precision=precision, **params)
    c = extent / 2.0
    sigma = extent / 10.0

    def bed(x, y):
        r2 = (x - c) ** 2 + (y - c) ** 2
        return x / 1000.0 + 0.5 * np.exp(-r2 / (2.0 * sigma * sigma))

    def stage(x, y):
        return np.maximum(np.where(x < c, dam_height, 0.5), bed(x, y))

    d.set_quantity("elevation", bed)
    d.set_quantity("stage", stage)
    d.set_quantity("friction", manning, location="centroids")
    d.set_boundary({"left": Dirichlet_boundary([dam_height, inflow, 0.0]), "right": Transmissive_boundary(d),
                    "bottom": Reflective_boundary(d), "top": Reflective_boundary(d)})
    return d


def ensemble_scenarios(count=64, seed=4242, triangle_range=(20000, 500000), precision="fp64", **params):
    """The batched-ensemble workload: ``count`` members with triangle counts
    log-uniform in ``triangle_range``, Manning n ~ U[0.01, 0.06], dam height
    ~ U[1, 4] m and inflow ~ U[0, 1] m^2/s (all seeded).  Returns the list of
    domains and the list of their parameters."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(triangle_range[0]), np.log(triangle_range[1])
    specs = []
    for i in range(count):
        specs.append(dict(triangles=int(np.exp(rng.uniform(lo, hi))), seed=int(seed * 1000 + i),
                          manning=float(rng.uniform(0.01, 0.06)), dam_height=float(rng.uniform(1.0, 4.0)),
                          inflow=float(rng.uniform(0.0, 1.0))))
    return [ensemble_member(precision=precision, **s, **params) for s in specs], specs

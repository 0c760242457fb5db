"""NairnMPM XML inputs of the benchmark workloads, written as a user of the reference would.

Units follow the reference's Legacy input units: lengths in mm, velocities in mm/s, moduli and
yield stresses in MPa, densities in g/cm^3, times in s (MPMReadHandler.cpp:1163-1170,
Common/Materials/IsotropicMat.cpp:40, MaterialBaseMPM.cpp:56-58).
"""
from __future__ import annotations

import math

_MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) -> uniform doubles in [0, 1) from the top 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform_array(self, n: int):
        """the next n uniform() values as a numpy array (same stream, vectorised)"""
        import numpy as np
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + np.uint64(0x9E3779B97F4A7C15) * np.arange(1, n + 1, dtype=np.uint64)
            self.state = (self.state + n * 0x9E3779B97F4A7C15) & _MASK
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def _header(analysis: int, cfl: float, max_time: float, extra: str = "", shape: str = "uGIMP") -> str:
    return f"""<?xml version='1.0'?>
<JANFEAInput version='3'>
  <Header>
    <Description>benchmark workload</Description>
    <Analysis>{analysis}</Analysis>
  </Header>
  <MPMHeader>
    <Timing CFL="{cfl!r}" max="{max_time!r}"/>
    <GIMP type="{shape}"/>
{extra}  </MPMHeader>
"""


def elastic_block_2d(cells: int = 64, block: int = 32, seed: int = 20260101, E: float = 1000.,
                     nu: float = 0.33, rho: float = 1.0, vmax: float = 10000., cfl: float = 0.2,
                     jitter: float = 0.05, steps_hint: int = 1000, material_xml: str | None = None,
                     shape: str = "uGIMP") -> str:
    """Plain 2D plane-strain elastic block: cells x cells grid of 1 mm cells, 2x2 particles per
    cell in the central block x block cells, each coordinate jittered by U(-jitter, jitter) mm
    with a SplitMix64 stream, v_x = vmax*sin(pi*(y-y0)/L), v_y = 0 (mm/s)."""
    rng = SplitMix64(seed)
    x0 = (cells - block) / 2.0
    pts = []
    for j in range(block):
        for i in range(block):
            for b in range(2):
                for a in range(2):
                    x = x0 + i + 0.25 + 0.5 * a + jitter * (2.0 * rng.uniform() - 1.0)
                    y = x0 + j + 0.25 + 0.5 * b + jitter * (2.0 * rng.uniform() - 1.0)
                    vx = vmax * math.sin(math.pi * (y - x0) / block)
                    pts.append(f'    <mp matl="1" lpx="0.5" lpy="0.5"><pt x="{x!r}" y="{y!r}"/>'
                               f'<vel x="{vx!r}" y="0.0"/></mp>')
    mat = material_xml or f"""  <Material Type="1" Name="block">
    <rho>{rho!r}</rho>
    <E>{E!r}</E>
    <nu>{nu!r}</nu>
  </Material>
"""
    return (_header(10, cfl, 1.0, shape=shape) +
            f"""  <Mesh output="file">
    <Grid xmin="0" xmax="{cells}" ymin="0" ymax="{cells}">
      <Horiz cellsize="1"/>
      <Vert cellsize="1"/>
    </Grid>
  </Mesh>
  <MaterialPoints>
    <PointList>
""" + "\n".join(pts) + """
    </PointList>
  </MaterialPoints>
""" + mat + "</JANFEAInput>\n")


def j2_block_3d(cells: int = 128, lo: int = 16, hi: int = 112, vx: float = 50000., E: float = 70000.,
                nu: float = 0.33, yield_: float = 300., Ep: float = 1000., rho: float = 2.7,
                cfl: float = 0.5, ppe: int = 8, shape: str = "uGIMP") -> str:
    """3D elastic-plastic block: cells^3 grid of 1 mm cells; a box [lo, hi]^3 filled with ppe
    particles per cell of IsoPlasticity (J2, linear hardening); the half with x < mid moves at
    +vx, the other half at -vx (mm/s): two halves colliding."""
    mid = (lo + hi) / 2.0
    return (_header(12, cfl, 1.0, f"    <MatlPtsPerElement>{ppe}</MatlPtsPerElement>\n", shape=shape) +
            f"""  <Mesh output="file">
    <Grid xmin="0" xmax="{cells}" ymin="0" ymax="{cells}" zmin="0" zmax="{cells}">
      <Horiz cellsize="1"/>
      <Vert cellsize="1"/>
      <Depth cellsize="1"/>
    </Grid>
  </Mesh>
  <MaterialPoints>
    <Body mat="1" vx="{vx!r}" vy="0" vz="0">
      <Box xmin="{lo}" xmax="{mid!r}" ymin="{lo}" ymax="{hi}" zmin="{lo}" zmax="{hi}"/>
    </Body>
    <Body mat="1" vx="{-vx!r}" vy="0" vz="0">
      <Box xmin="{mid!r}" xmax="{hi}" ymin="{lo}" ymax="{hi}" zmin="{lo}" zmax="{hi}"/>
    </Body>
  </MaterialPoints>
  <Material Type="9" Name="J2 block">
    <rho>{rho!r}</rho>
    <E>{E!r}</E>
    <nu>{nu!r}</nu>
    <yield>{yield_!r}</yield>
    <Ep>{Ep!r}</Ep>
  </Material>
</JANFEAInput>
""")


def grid_bcs(xml: str, shapes, block: str = "GridBCs", child: str = "DisBC") -> str:
    """xml with a <GridBCs> block added: shapes is a list of (tag, attributes, [DisBC attributes]),
    tag "BCLine" (2D) or "BCBox" (3D), e.g.
    ("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [{"dir": 2, "vel": 0}]); a string in the
    list is an empty element of that name (<net/>, <perParticle/>)"""
    def attrs(d):
        return " ".join(f'{k}="{v!r}"' if isinstance(v, float) else f'{k}="{v}"' for k, v in d.items())
    body = f"  <{block}>\n"
    for tag, a, bcs in shapes:
        body += f"    <{tag} {attrs(a)}>\n"
        for b in bcs:
            body += f"      <{b}/>\n" if isinstance(b, str) else f"      <{child} {attrs(b)}/>\n"
        body += f"    </{tag}>\n"
    body += f"  </{block}>\n"
    return xml.replace("</JANFEAInput>", body + "</JANFEAInput>", 1)


def particle_bcs(xml: str, shapes) -> str:
    """xml with a <ParticleBCs> block of <LoadBC> loads, shapes as for grid_bcs, e.g.
    ("BCLine", {...}, ["net", {"dir": 1, "load": 2.0, "style": 1}])"""
    return grid_bcs(xml, shapes, block="ParticleBCs", child="LoadBC")


def traction_bcs(xml: str, shapes) -> str:
    """xml with a <ParticleBCs> block of <TractionBC> conditions, shapes as for grid_bcs, e.g.
    ("BCLine", {...}, [{"dir": 2, "face": 3, "stress": 1.5, "style": 1}])"""
    return grid_bcs(xml, shapes, block="ParticleBCs", child="TractionBC")


def periodic_xpic(xml: str, order: int = 2, fmpm: bool = False, steps: int | None = 1,
                  time_ms: float | None = None, cfl: float | None = None) -> str:
    """xml with a PeriodicXPIC custom task: XPIC(order) or FMPM(order) every `steps` steps, or by
    time (periodicTime in ms) or by a CFL factor"""
    params = [("FMPMOrder" if fmpm else "XPICOrder", order)]
    if steps is not None:
        params.append(("periodicSteps", steps))
    if time_ms is not None:
        params.append(("periodicTime", time_ms))
    if cfl is not None:
        params.append(("periodicCFL", cfl))
    body = "  <CustomTasks>\n    <Schedule name=\"PeriodicXPIC\">\n"
    for k, v in params:
        body += f'      <Parameter name="{k}">{v!r}</Parameter>\n'
    body += "    </Schedule>\n  </CustomTasks>\n"
    return xml.replace("</JANFEAInput>", body + "</JANFEAInput>", 1)


def mooney_material_xml(G1: float = 8., G2: float = 2., K: float = 50., rho: float = 1.1,
                        name: str = "mooney") -> str:
    """<Material Type="8"> (Mooney-Rivlin) with G1, G2 and K in MPa."""
    return f"""  <Material Type="8" Name="{name}">
    <rho>{rho!r}</rho>
    <G1>{G1!r}</G1>
    <G2>{G2!r}</G2>
    <K>{K!r}</K>
  </Material>
"""


def neohookean_block_3d(cells: int = 32, lo: int = 8, hi: int = 24, vx: float = 20000., G: float = 10.,
                        K: float = 50., rho: float = 1.1, cfl: float = 0.5, material_xml: str | None = None,
                        shape: str = "uGIMP") -> str:
    """3D neo-Hookean block (C3/C5 material) with two colliding halves; material_xml replaces the
    material (e.g. mooney_material_xml())."""
    mid = (lo + hi) / 2.0
    mat = material_xml or f"""  <Material Type="28" Name="rubber">
    <rho>{rho!r}</rho>
    <G>{G!r}</G>
    <K>{K!r}</K>
  </Material>
"""
    return (_header(12, cfl, 1.0, "    <MatlPtsPerElement>8</MatlPtsPerElement>\n", shape=shape) +
            f"""  <Mesh output="file">
    <Grid xmin="0" xmax="{cells}" ymin="0" ymax="{cells}" zmin="0" zmax="{cells}">
      <Horiz cellsize="1"/>
      <Vert cellsize="1"/>
      <Depth cellsize="1"/>
    </Grid>
  </Mesh>
  <MaterialPoints>
    <Body mat="1" vx="{vx!r}" vy="0" vz="0">
      <Box xmin="{lo}" xmax="{mid!r}" ymin="{lo}" ymax="{hi}" zmin="{lo}" zmax="{hi}"/>
    </Body>
    <Body mat="1" vx="{-vx!r}" vy="0" vz="0">
      <Box xmin="{mid!r}" xmax="{hi}" ymin="{lo}" ymax="{hi}" zmin="{lo}" zmax="{hi}"/>
    </Body>
  </MaterialPoints>
""" + mat + """</JANFEAInput>
""")


def j2_block_2d(cells: int = 48, block: int = 24, seed: int = 7, E: float = 2000., nu: float = 0.3,
                yield_: float = 10., Ep: float = 20., rho: float = 1.0, vmax: float = 20000.,
                cfl: float = 0.5) -> str:
    """2D plane-strain J2 block (the C4 sweep's plastic member), C1 layout and velocity profile."""
    mat = f"""  <Material Type="9" Name="J2 block">
    <rho>{rho!r}</rho>
    <E>{E!r}</E>
    <nu>{nu!r}</nu>
    <yield>{yield_!r}</yield>
    <Ep>{Ep!r}</Ep>
  </Material>
"""
    return elastic_block_2d(cells=cells, block=block, seed=seed, vmax=vmax, cfl=cfl, material_xml=mat)


def contact_cylinders_3d(cells: int = 160, radius: float = 20., length: float = 60., gap: float = 4.,
                         vx: float = 10000., G: float = 10., K: float = 50., rho: float = 1.1,
                         friction: float = 0.3, shape: str = "lCPDI", normals: int = 2,
                         cfl: float = 0.5, law_type: int = 61) -> str:
    """Two neo-Hookean cylinders (axes along z) closing at 2*vx mm/s, in multimaterial mode with a
    Coulomb friction law: the contact workload. The cylinders are two materials with the same
    properties so that each has its own velocity field; the third material is the contact law."""
    c = cells / 2.0
    xa, xb = c - radius - gap / 2.0, c + radius + gap / 2.0
    z0, z1 = c - length / 2.0, c + length / 2.0
    header = (f"<?xml version='1.0'?>\n<JANFEAInput version='3'>\n  <Header>\n"
              f"    <Description>contact workload</Description>\n    <Analysis>12</Analysis>\n  </Header>\n"
              f"  <MPMHeader>\n    <MatlPtsPerElement>8</MatlPtsPerElement>\n"
              f"    <Timing CFL=\"{cfl!r}\" max=\"1.0\"/>\n    <GIMP type=\"{shape}\"/>\n"
              f"    <MultiMaterialMode Normals=\"{normals}\">\n      <Friction law=\"3\"/>\n"
              f"    </MultiMaterialMode>\n  </MPMHeader>\n")
    mats = ""
    for k in (1, 2):
        mats += f"""  <Material Type="28" Name="cylinder {k}">
    <rho>{rho!r}</rho>
    <G>{G!r}</G>
    <K>{K!r}</K>
  </Material>
"""
    mats += f"""  <Material Type="{law_type}" Name="friction">
    <coeff>{friction!r}</coeff>
  </Material>
"""
    return (header + f"""  <Mesh output="file">
    <Grid xmin="0" xmax="{cells}" ymin="0" ymax="{cells}" zmin="0" zmax="{cells}">
      <Horiz cellsize="1"/>
      <Vert cellsize="1"/>
      <Depth cellsize="1"/>
    </Grid>
  </Mesh>
  <MaterialPoints>
    <Body mat="1" vx="{vx!r}" vy="0" vz="0">
      <Cylinder axis="z" xmin="{xa - radius!r}" xmax="{xa + radius!r}" ymin="{c - radius!r}" ymax="{c + radius!r}" zmin="{z0!r}" zmax="{z1!r}"/>
    </Body>
    <Body mat="2" vx="{-vx!r}" vy="0" vz="0">
      <Cylinder axis="z" xmin="{xb - radius!r}" xmax="{xb + radius!r}" ymin="{c - radius!r}" ymax="{c + radius!r}" zmin="{z0!r}" zmax="{z1!r}"/>
    </Body>
  </MaterialPoints>
""" + mats + "</JANFEAInput>\n")


def contact_blocks_2d(cells: int = 32, w: int = 8, h: int = 8, gap: float = 1., vx: float = 5000.,
                      vy: float = 0., E: float = 500., nu: float = 0.3, rho: float = 1.0, friction: float = 0.3,
                      normals: int = 2, law_type: int = 61, cfl: float = 0.5, offset_y: float = 0.) -> str:
    """Two plane-strain elastic blocks (materials 1 and 2) closing along x at 2*vx mm/s, the right
    one sliding at vy, with the contact law as material 3 (the 2D multimaterial case)."""
    c = cells / 2.0
    xa0, xa1 = c - gap / 2.0 - w, c - gap / 2.0
    xb0, xb1 = c + gap / 2.0, c + gap / 2.0 + w
    y0, y1 = c - h / 2.0, c + h / 2.0
    mats = ""
    for k in (1, 2):
        mats += f"""  <Material Type="1" Name="block {k}">
    <rho>{rho!r}</rho>
    <E>{E!r}</E>
    <nu>{nu!r}</nu>
  </Material>
"""
    mats += f"""  <Material Type="{law_type}" Name="friction">
    <coeff>{friction!r}</coeff>
  </Material>
"""
    return (f"""<?xml version='1.0'?>
<JANFEAInput version='3'>
  <Header>
    <Description>2D contact</Description>
    <Analysis>10</Analysis>
  </Header>
  <MPMHeader>
    <MatlPtsPerElement>4</MatlPtsPerElement>
    <Timing CFL="{cfl!r}" max="1.0"/>
    <GIMP type="uGIMP"/>
    <MultiMaterialMode Normals="{normals}">
      <Friction law="3"/>
    </MultiMaterialMode>
  </MPMHeader>
  <Mesh output="file">
    <Grid xmin="0" xmax="{cells}" ymin="0" ymax="{cells}">
      <Horiz cellsize="1"/>
      <Vert cellsize="1"/>
    </Grid>
  </Mesh>
  <MaterialPoints>
    <Body mat="1" vx="{vx!r}" vy="0">
      <Rect xmin="{xa0!r}" xmax="{xa1!r}" ymin="{y0!r}" ymax="{y1!r}"/>
    </Body>
    <Body mat="2" vx="{-vx!r}" vy="{vy!r}">
      <Rect xmin="{xb0!r}" xmax="{xb1!r}" ymin="{y0 + offset_y!r}" ymax="{y1 + offset_y!r}"/>
    </Body>
  </MaterialPoints>
""" + mats + "</JANFEAInput>\n")


def sweep_instance_2d(k: int, pad: int = 8, cfl: float = 0.2):
    """Member k of the batched 2D sweep: seed 1000+k; Np = round(exp(U(ln 2000, ln 200000)))
    particles as a square block at 4 per cell with `pad` cells of grid around it; elastic or J2
    with probability 1/2; E ~ U(500, 5000) MPa, nu ~ U(0.25, 0.40), yield ~ U(5, 50) MPa,
    Ep = 0.01 E, rho = 1 g/cm^3. Returns (xml, particle_fix, is_j2): particle_fix(d) applies the
    plain 2D block's velocity profile and +-0.05 mm jitter to the particle fields d in place."""
    rng = SplitMix64(1000 + k)
    npart = round(math.exp(math.log(2000.) + rng.uniform() * (math.log(200000.) - math.log(2000.))))
    side = max(2, round(math.sqrt(npart / 4.)))
    cells = side + 2 * pad
    E = 500. + 4500. * rng.uniform()
    nu = 0.25 + 0.15 * rng.uniform()
    j2 = rng.uniform() < 0.5
    yld = 5. + 45. * rng.uniform()
    if j2:
        mat = f"""  <Material Type="9" Name="J2 {k}">
    <rho>1.0</rho>
    <E>{E!r}</E>
    <nu>{nu!r}</nu>
    <yield>{yld!r}</yield>
    <Ep>{0.01 * E!r}</Ep>
  </Material>
"""
    else:
        mat = f"""  <Material Type="1" Name="elastic {k}">
    <rho>1.0</rho>
    <E>{E!r}</E>
    <nu>{nu!r}</nu>
  </Material>
"""
    xml = (_header(10, cfl, 1.0, "    <MatlPtsPerElement>4</MatlPtsPerElement>\n") +
           f"""  <Mesh output="file">
    <Grid xmin="0" xmax="{cells}" ymin="0" ymax="{cells}">
      <Horiz cellsize="1"/>
      <Vert cellsize="1"/>
    </Grid>
  </Mesh>
  <MaterialPoints>
    <Body mat="1" vx="0" vy="0">
      <Rect xmin="{pad}" xmax="{pad + side}" ymin="{pad}" ymax="{pad + side}"/>
    </Body>
  </MaterialPoints>
""" + mat + "</JANFEAInput>\n")

    def particle_fix(d):
        """C1 velocity profile v_x = 10 m/s sin(pi (y - y0)/L) and +-0.05 mm position jitter"""
        import numpy as np
        n = d.shape[1]
        u = rng.uniform_array(2 * n).reshape(n, 2)
        d[F_POS_X] += 0.05 * (2.0 * u[:, 0] - 1.0)
        d[F_POS_Y] += 0.05 * (2.0 * u[:, 1] - 1.0)
        d[F_VEL_X] = 10000. * np.sin(np.pi * (d[F_POS_Y] - pad) / side)
        d[F_ORIG_X] = d[F_POS_X]
        d[F_ORIG_Y] = d[F_POS_Y]

    return xml, particle_fix, j2


def serve_request_3d(k: int, pad: int = 8, nmin: int = 50000, nmax: int = 2000000,
                     smin: int = 200, smax: int = 2000):
    """Request k of the served 3D stream: seed 5000+k; Np = round(exp(U(ln nmin, ln nmax)))
    particles as a cube block at 2x2x2 per 1 mm cell with `pad` cells of grid around it; J2 (the
    7M-particle block's material, halves colliding at +-50 m/s) or neo-Hookean (G 10 MPa, K 50 MPa,
    rho 1.1, halves at +-5 m/s) with probability 1/2; steps uniform in [smin, smax].
    Returns (xml, steps, is_j2)."""
    rng = SplitMix64(5000 + k)
    npart = math.exp(math.log(nmin) + rng.uniform() * (math.log(nmax) - math.log(nmin)))
    side = max(2, round((npart / 8.) ** (1. / 3.)))
    j2 = rng.uniform() < 0.5
    steps = smin + int(rng.uniform() * (smax - smin + 1))
    cells = side + 2 * pad
    if j2:
        xml = j2_block_3d(cells=cells, lo=pad, hi=pad + side)
    else:
        xml = neohookean_block_3d(cells=cells, lo=pad, hi=pad + side, vx=5000.)
    return xml, steps, j2


# rows of the particle field array (nairnmpm.FIELDS)
F_POS_X, F_POS_Y, F_VEL_X, F_ORIG_X, F_ORIG_Y = 0, 1, 3, 34, 35

"""Drop-in locate_particles (locate_particles.py:12-66) on the GPU."""
from .pic_functions import _locate_annulus, _locate_box


def locate_particles(geometry, nparticle, swarm_active, swarm_x, swarm_z, swarm_rad, swarm_theta,
                     hx, hz, hrad, htheta, x_V, z_V, rad_V, theta_V, icon_V, nelx, Rinner):
    """locate_particles___box (box) or ___annulus (quarter, half, eighth) plus
    the reference's range checks (locate_particles.py:61-64), which raise
    SystemExit like its exit(); the annulus itself exits as the reference."""
    if geometry == "box":
        r, t, iel, flags = _locate_box(nparticle, swarm_active, swarm_x, swarm_z, hx, hz, x_V,
                                       z_V, icon_V, nelx)
    elif geometry in ("quarter", "half", "eighth"):
        r, t, iel, flags = _locate_annulus(nparticle, swarm_active, swarm_rad, swarm_theta, hrad,
                                           htheta, rad_V, theta_V, icon_V, nelx, Rinner)
    elif geometry == "annulus":
        raise SystemExit("locate particles not implemented for annulus geometry")
    else:
        raise SystemExit("locate particles not implemented for this geometry")
    if flags[0] or flags[1]:
        raise SystemExit("r value out of bounds")
    if flags[2]:
        raise SystemExit("t value out of bounds")
    return r, t, iel

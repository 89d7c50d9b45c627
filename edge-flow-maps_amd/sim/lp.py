"""Drop-in for the reference's sim/lp.py: the laden-particle kernels.

Same function names, argument order and meaning as the reference
(sim/lp.py:6-218); the work runs in libedge_hip.so.  Particle fields are
`ti.Vector.field(3, float, shape=max_num)` Fields (3 planes of max_num on the
GPU); only particles I < laden_particle_num are live.  Grid fields are the
three MAC face grids.
"""
from edge_hip import _lib
from edge_hip.field import as_ptr, stream_handle, touch


def _cells(u_x):
    return u_x.shape[0] - 1, u_x.shape[1], u_x.shape[2]


def _n(f):
    return f.shape[0]


def update_particles_acc(u_x, u_y, u_z, laden_particles_vel, laden_particles_pos, laden_particles_acc, dx, curr_dt,
                         laden_particle_num, St, gravity):
    """sim/lp.py:6-29."""
    _lib.call("edge_lp_acc", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), *_cells(u_x), as_ptr(laden_particles_vel),
              as_ptr(laden_particles_pos), as_ptr(laden_particles_acc), _n(laden_particles_acc),
              int(laden_particle_num), float(dx), float(curr_dt), float(St), float(gravity), stream_handle())
    touch(laden_particles_acc)


def update_particles_velocity(laden_particles_acc, laden_particles_vel, laden_particles_pos, curr_dt, dx,
                              laden_particle_num, domain_range_x, domain_range_y, domain_range_z):
    """sim/lp.py:32-70."""
    _lib.call("edge_lp_vel", as_ptr(laden_particles_acc), as_ptr(laden_particles_vel), as_ptr(laden_particles_pos),
              _n(laden_particles_vel), int(laden_particle_num), float(curr_dt), float(dx), float(domain_range_x),
              float(domain_range_y), float(domain_range_z), stream_handle())
    touch(laden_particles_vel)


def calculate_velocity_diff(u_x, u_y, u_z, laden_particles_vel, laden_particles_vel_diff, laden_particles_pos, dx,
                            laden_particle_num, St, density_ratio, epiral_radius_ratio):
    """sim/lp.py:73-97."""
    _lib.call("edge_lp_vel_diff", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), *_cells(u_x), as_ptr(laden_particles_vel),
              as_ptr(laden_particles_vel_diff), as_ptr(laden_particles_pos), _n(laden_particles_vel),
              int(laden_particle_num), float(dx), float(St), float(density_ratio), float(epiral_radius_ratio),
              stream_handle())
    touch(laden_particles_vel_diff)


def laden_particles_back_with_computing_particle(drag_x, drag_y, drag_z, u_x, u_y, u_z, laden_particles_vel,
                                                 laden_particles_vel_diff, laden_particles_pos,
                                                 laden_particle_ref_radius, dx, laden_particle_num, laden_radius,
                                                 domain_range_x, domain_range_y, domain_range_z, res_x, res_y, res_z):
    """sim/lp.py:100-156 (u_* and the velocities are unused there too)."""
    _lib.call("edge_lp_p2g", as_ptr(drag_x), as_ptr(drag_y), as_ptr(drag_z), as_ptr(laden_particles_vel_diff),
              as_ptr(laden_particles_pos), as_ptr(laden_particle_ref_radius), _n(laden_particles_pos),
              int(laden_particle_num), float(dx), float(laden_radius), float(domain_range_x), float(domain_range_y),
              float(domain_range_z), int(res_x), int(res_y), int(res_z), stream_handle())
    touch(drag_x, drag_y, drag_z)


def update_particles_position(laden_particles_vel, laden_particles_pos, curr_dt, laden_particle_num, domain_range_x,
                              domain_range_y, domain_range_z):
    """sim/lp.py:158-181."""
    _lib.call("edge_lp_pos", as_ptr(laden_particles_vel), as_ptr(laden_particles_pos), _n(laden_particles_pos),
              int(laden_particle_num), float(curr_dt), float(domain_range_x), float(domain_range_y),
              float(domain_range_z), stream_handle())
    touch(laden_particles_pos)


def calculate_viscous_force(u_x, u_y, u_z, vis_x, vis_y, vis_z, dx, Re):
    """sim/lp.py:183-218."""
    _lib.call("edge_lp_viscous", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), as_ptr(vis_x), as_ptr(vis_y), as_ptr(vis_z),
              *_cells(u_x), float(dx), float(Re), stream_handle())
    touch(vis_x, vis_y, vis_z)


def advect_dye_particles(u_x, u_y, u_z, dye_pos, dx, curr_dt, dye_num):
    """Passive dye (no reference counterpart; the 512x256x256 torus case's dye):
    each live particle q < dye_num takes the RK4 step of a forward face-map
    point (run_edge.py:336-357, +dt, the B-spline MAC velocity at every stage)
    and is clamped into the domain (as sim/lp.py:158-181 clamps positions)."""
    _lib.call("edge_dye_advance", as_ptr(u_x), as_ptr(u_y), as_ptr(u_z), as_ptr(dye_pos), _n(dye_pos), int(dye_num),
              *_cells(u_x), float(dx), float(curr_dt), stream_handle())
    touch(dye_pos)

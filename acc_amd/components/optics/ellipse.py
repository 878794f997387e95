"""Elliptic segment files for the tapered guide (the role of
acc/test/ellipse.py:100-165, create_guide_tapering_data).

Rows are `h1 h2 w1 w2` (full heights/widths at the entrance and exit of each
segment) under three comment lines starting with "c".  The tapered-guide
loader drops the last data row (acc/components/optics/guide_tapering.py:108-111),
so `extra_row=True` appends one more segment to keep `segments` usable ones.
"""
import numpy as np


def ellipse_half_width(z, a, b):
    """Half width of an ellipse with semi-axes a (across) and b (along z), at
    distance z from its vertex: a sqrt(z/b (2 - z/b))."""
    return a * np.sqrt(z / b * (2 - z / b))


def create_guide_tapering_data(path, z0, z_start, z_exit, segments, major_axis_len, minor_axis_len,
                               extra_row=False):
    """Write a segment file for an elliptic guide section.

    z0: z of the ellipse centre; z_start / z_exit: guide openings (same frame);
    segments: number of segments between them; major/minor_axis_len: full axis
    lengths of the ellipse.
    """
    b = major_axis_len / 2.
    a = minor_axis_len / 2.
    nrows = segments + (1 if extra_row else 0)
    interval = (z_exit - z_start) / segments
    zs = z_start + interval * np.arange(nrows) - (z0 - b)
    lines = [
        'c use with Guide_tapering.comp',
        'c start position: {} from focal point, length: {}, number of segments: {}'.format(
            z_start, z_exit - z_start, segments),
        'c h1(i)     h2(i)   w1(i)    w2(i)',
    ]
    for z in zs:
        h1 = 2 * ellipse_half_width(z, a, b)
        h2 = 2 * ellipse_half_width(z + interval, a, b)
        lines.append('{} {} {} {}'.format(h1, h2, h1, h2))
    with open(path, 'wt') as stream:
        for line in lines:
            stream.write(line + '\n')
    return path


def config2_taper_file(path):
    """100 usable segments over 10 m narrowing 6 -> 3 cm: an ellipse centred at
    the entrance with semi-minor axis 3 cm and semi-major axis 10/sqrt(3/4) m,
    so the full width is 0.06 sqrt(1 - (z/11.547)^2) (0.03 m at z = 10 m)."""
    b = 10.0 / np.sqrt(0.75)
    return create_guide_tapering_data(path, z0=0.0, z_start=0.0, z_exit=10.0, segments=100,
                                      major_axis_len=2 * b, minor_axis_len=0.06, extra_row=True)

"""awkward_amd -- an MI355X-native ``hip`` backend for Awkward Array's
jagged-array kernel layer.

The kernels are hand-written HIP for gfx950 in ``libawkward_hip.so``
(``awkward_amd/csrc``), served under the reference's kernel names through
``HipBackend`` (``backend[(name, *dtypes)](*args)``), and driven by a host
layer that issues the reference's call sequences for ``sum``, ``prod``,
``any``, ``all``, ``max``, ``min``, ``argmax``, ``argmin``, ``count``,
``num``, ``combinations``,
``argcombinations``,
jagged and range ``__getitem__``, ``sort`` and ``argsort`` at axis=-1.
Buffers enter and leave without copies through DLPack (``from_dlpack``)
and the Form/length/container triple (``to_buffers`` / ``from_buffers``).
"""

from awkward_amd.highlevel import (  # noqa: F401
    Array,
    all,
    any,
    argcombinations,
    argmax,
    argmin,
    argsort,
    combinations,
    count,
    from_list,
    from_offsets,
    local_index,
    max,
    min,
    num,
    prod,
    sort,
    sum,
    to_backend,
)
from awkward_amd.interop import DLPackDevice, from_buffers, from_dlpack, to_buffers  # noqa: F401

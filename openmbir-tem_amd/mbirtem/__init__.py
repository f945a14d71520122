"""mbirtem: Python view of the MI355X-native MBIR reconstruction engine.

The engine is libmbir_tem.so (HIP kernels + C-ABI, include/mbir_tem.h); this
package only binds it (capi) and provides seeded synthetic workloads (phantom,
workloads).
"""
from . import capi, phantom  # noqa: F401

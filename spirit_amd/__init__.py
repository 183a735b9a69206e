"""spirit_amd: the HIP (MI355X, gfx950) engine for Spirit's effective-field and
LLG/GNEB step path, with the reference's ctypes Python interface.

Modules mirror the reference package ``spirit`` (core/python/spirit): ``state``,
``simulation``, ``system``, ``configuration``, ``hamiltonian``, ``geometry``,
``parameters.llg``, ``parameters.gneb``, ``chain``, ``transition``,
``quantities``, ``constants``, ``log``; ``hip`` holds the engine-only
extensions.  ``libSpirit.so`` is loaded from this directory; build it with
``python -m spirit_amd.build``.
"""
__version__ = "2.2.0+hip"

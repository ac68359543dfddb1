"""Plugin protocol of the simulation (reference ``nofluidx3d/interactions/Interaction.py:15-40``).

A plugin is registered with ``Simulation.register``, which calls ``register(simulation)``: the
plugin allocates its device state (``createBuffers``), binds its kernel (``build``, which sets
``scope`` and ``knl``) and hands the kernel its arguments (``setArgs``).  Each reference time step
then calls ``beforeTimeStep(globalTime)`` on every plugin and ``enqueue()`` on every plugin, in
registration order.

The built-in plugins bind ``knl`` to a :class:`DeviceCall` of a libnfx entry point (precompiled
for gfx950); third-party plugins may still go through ``KernelBuilder``.

Two hooks belong to this implementation:
- ``contribute(plan)``: a plugin that can describe its term to the fused device step
  (``nfx_run_steps``) fills its fields of the ``nfx_plan`` and returns True;
- ``state_form(fused)`` / ``sync_forms()`` / ``forms_restored()``: plugins whose state the fused
  step keeps in another representation (the viscoplastic law's plastic deviation) convert between
  the two when the stepping path changes.
"""

from __future__ import annotations


class DeviceCall:
    """A libnfx entry point bound to an argument provider: ``call()`` enqueues it on the active
    context's stream with the arguments ``args()`` returns at that moment (device pointers are read
    when the call is made, so buffers swapped after ``build`` are picked up)."""

    def __init__(self, entry: str, args):
        self.entry = entry
        self.args = args

    def __call__(self):
        from nofluidx3d_amd import _lib, openCL

        _lib.call(self.entry, openCL.handle(), *self.args())


class Interaction:
    def __init__(self, scope=None):
        self.scope = scope
        self.knl = None

    def register(self, simulation):
        self.simulation = simulation
        self.createBuffers()
        self.build()
        self.setArgs()

    def createBuffers(self):
        pass

    def build(self):
        pass

    def setArgs(self):
        pass

    def beforeTimeStep(self, globalTime):
        pass

    def enqueue(self):
        self.knl()

    def contribute(self, plan) -> bool:
        return False

    def state_form(self, fused: bool):
        """Called before steps; ``fused`` is True when the fused device step will advance the
        plugin's state next."""

    def sync_forms(self):
        """Bring every state representation up to date (before a snapshot of the state)."""

    def forms_restored(self):
        """Every state representation was just restored from a snapshot taken after sync_forms."""

    def state_buffers(self):
        """Device tensors holding this plugin's evolving state (for snapshots)."""
        return []

"""Plugin base class (reference ``nofluidx3d/interactions/Interaction.py:15-40``).

Protocol: ``register(simulation)`` calls ``createBuffers()``, ``build()``, ``setArgs()``;
``Simulation.timeStep`` calls every plugin's ``beforeTimeStep(globalTime)`` and then every
plugin's ``enqueue()`` in registration order.  ``enqueue`` launches the plugin's precompiled
kernel on the context's stream.

``contribute(plan)`` is this implementation's fusion hook: a plugin that can describe its term to
the fused step (``nfx_run_steps``) fills its fields of the ``nfx_plan`` and returns True.
"""


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

    def state_form(self, shape: bool):
        """Called before steps: ``shape`` is True when the tiled fused step will advance the
        plugin's state (plugins with a second state representation convert here)."""

    def sync_forms(self):
        """Bring every state representation up to date (before a snapshot of the state)."""

    def forms_restored(self):
        """Every state representation was just restored from a snapshot taken after sync_forms."""

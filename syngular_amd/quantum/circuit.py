"""Gate sequence over a Qbit register (drop-in for
syngular.quantum.Circuit, quantum/circuit.py:7-46)."""
from __future__ import annotations

from .qbit import Qbit


class Circuit:
    def __init__(self, size, bond=2, initializer="ground", structure=[]):
        # quantum/circuit.py:9-19.  As in the reference the default
        # ``structure`` list is one object shared by every Circuit built
        # without one, so ``add`` on such a circuit extends it for all.
        self.initializer = initializer
        self.size = size
        self.structure = structure
        self.current_step = 0
        self.current_state = None
        self.states = []
        self.reset()

    def run(self):
        for _ in range(self.current_step, len(self.structure)):
            self.step()

    def reset(self):
        if self.initializer == "ground":
            self.current_state = Qbit(self.size)
            self.states.append(self.current_state)

    def step(self):
        self.current_state @= self.structure[self.current_step]
        self.states.append(self.current_state)
        self.current_step += 1

    def add(self, gate):
        self.structure.append(gate)

    def get(self, index=-1):
        return self.states[index]

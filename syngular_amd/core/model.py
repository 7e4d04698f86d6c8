"""Tensor-train layers (drop-in for syngular.core, core/model.py:7-226).

``Linear`` holds its weight as a MatrixProductOperator and its forward is
``weight @ inputs`` (core/model.py:201-206): the MPO x MPS apply kernel on
every site, then the reference's ``>> min_bond`` QR truncation, all on the
device.  Extension: a Linear called on a list of states runs the whole batch
in grouped launches (``matmul_batch``), with per-state results equal to
``weight @ x`` one by one.  The reference's console output of ``train`` and
``describe`` is kept; ``train`` adds a random (2,2)x(2,2) bond-4 operator to
every weight as the reference does (core/model.py:33-42).
"""
from __future__ import annotations

import numpy as np
import torch

from .. import linalg, ops
from ..tensor import MatrixProductOperator, MatrixProductState


def matmul_batch(weight, states):
    """[weight @ x for x in states] (tensor/matrix_product_operator.py:175-192
    per state) in grouped launches: one MPO x MPS apply launch for every site
    of every state, then the QR truncation to each state's min bond
    (tensor/matrix_product_state.py:432-468), bond by bond, with one QR launch
    and two GEMM launches per bond over all the states that need it.  States
    whose site count differs from the operator's, or a non-default matmul /
    rounding mode, go through ``weight @ x`` one by one."""
    states = list(states)
    n = weight.sites_number
    if (MatrixProductOperator.MATMUL_MODE != "standard" or MatrixProductState.ROUNDING != "qr"
            or any(x.sites_number != n for x in states)):
        return [weight @ x for x in states]
    dims = [min(min(weight.bond_shape), min(x.bond_shape)) for x in states]
    flat = ops.apply_many([(weight.sites[k], x.sites[k]) for x in states for k in range(n)])
    outs = [MatrixProductState.from_sites(flat[i * n:(i + 1) * n]) for i in range(len(states))]
    if not MatrixProductOperator.COMPRESS:
        return outs
    todo = [i for i, o in enumerate(outs) if dims[i] < min(o.bond_shape)]
    work = {i: outs[i].copy() for i in todo}
    params = {i: 0 for i in todo}
    for k in range(n - 1):
        Ls = [linalg.left_unfold(work[i].sites[k], work[i].shape[k]) for i in todo]
        Rs = [linalg.right_unfold(work[i].sites[k + 1], work[i].shape[k + 1]) for i in todo]
        Qs = ops.qr_many([(L, dims[i]) for L, i in zip(Ls, todo)])
        Ss = [torch.empty((Q.shape[1], L.shape[1]), dtype=torch.float64, device=L.device) for Q, L in zip(Qs, Ls)]
        ops.gemm_many([(Q.t(), L, S, 1.0, 0.0) for Q, L, S in zip(Qs, Ls, Ss)])
        Ws = [torch.empty((S.shape[0], Rm.shape[1]), dtype=torch.float64, device=S.device) for S, Rm in zip(Ss, Rs)]
        ops.gemm_many([(S, Rm, W, 1.0, 0.0) for S, Rm, W in zip(Ss, Rs, Ws)])
        for i, Q, W in zip(todo, Qs, Ws):
            t = work[i]
            t.shape[k] = tuple(t.shape[k][:-1]) + (dims[i],)
            t.shape[k + 1] = (dims[i],) + tuple(t.shape[k + 1][1:])
            t.sites[k] = linalg.reshape(Q, t.shape[k])
            t.sites[k + 1] = linalg.reshape(W, t.shape[k + 1])
            params[i] += t.sites[k].numel()
    for i in todo:
        work[i].parameters_number = params[i] + work[i].sites[-1].numel()
        outs[i] = work[i]
    return outs


class Model:
    # core/model.py:7-82

    def __init__(self, layers):
        self.layers = layers

    def predict(self, inputs):
        values = inputs
        for layer in self.layers:
            values = layer(values)
        return values

    def feed_forward(self, sample):
        values = sample
        for layer in self.layers:
            values = layer(values)
        return values

    def build(self):
        for layer in self.layers:
            if not layer.is_built:
                layer.build(None)
                layer.is_built = True

    def train(self, x, y, batchsize=32, epochs=1, verbose=1):
        if verbose:
            print("[START] Training ")
        for e in range(epochs):
            if verbose:
                print(f"Epoch {str(e + 1)} : ")
            for layer in self.layers:
                for weight in layer.trainable_tensor_weights:
                    weight["weight"] += MatrixProductOperator.random((2, 2), (2, 2), (4,))
        if verbose:
            print("[END] Training ")

    def draw(self):
        return "".join(layer.draw() for layer in self.layers)

    def describe(self):
        strict_layer = ["output"]
        params = params_fictive = layers_num = 0
        rule = "-" * 111
        for layer in self.layers:
            info = layer.describe()
            layers_num += 1
            params += info["param"]
            params_fictive += info["fictive-param"]
            print(rule)
            if info["type"] not in strict_layer:
                saved = info["fictive-param"] - info["param"]
                print(f"|  Layer [{info['type']}] -> {info['name']} \t Parameters: {info['param']} \t "
                      f"Saved: {saved} (compression {round(saved / info['fictive-param'], 4) * 100}%)")
            else:
                print(f"|  Layer [{info['type']}] ->  {info['name']} \t ")
            print(rule)
        print("=" * 111)
        print(f"| \t Total number of layers : {layers_num}")
        print(f"| \t Total number of parameters : {params} (v.s {params_fictive})")
        print(f"| \t Total compression factor : {round((params_fictive - params) / params_fictive, 4) * 100}%")
        print("=" * 111)


class Layer:
    # core/model.py:84-160

    def __init__(self, name):
        self.name = name
        self.trainable_tensor_weights = []
        self.trainable_tensor_bias = []
        self.is_built = False

    def __call__(self, inputs):
        input_shape = inputs.shape if not isinstance(inputs, (list, tuple)) else None
        if not self.is_built:
            self.build(input_shape)
            self.is_built = True

    def build(self, input_shape):
        pass

    def draw(self):
        out = ""
        for weight in self.trainable_tensor_weights:
            mp = weight["weight"]
            out += "\t" + "|   " * mp.sites_number + "\n"
            out += "\t" + ("O---" * (mp.sites_number - 1)) + "O" + "\n"
            out += "\t" + "|   " * mp.sites_number + "\n"
        return out

    def describe(self):
        total = sum(w["weight"].parameters_number for w in self.trainable_tensor_weights)
        total += sum(b["bias"].parameters_number for b in self.trainable_tensor_bias)
        true = sum(w["weight"].real_parameters_number for w in self.trainable_tensor_weights)
        true += sum(b["bias"].real_parameters_number for b in self.trainable_tensor_bias)
        return {"param": total, "fictive-param": true, "name": self.name, "type": self.type}

    def add_weight(self, input_shape, output_shape, bond_shape, name=None, initializer="normal"):
        if name is None:
            name = f"weight_{np.random.randint(0, 999999)}"
        weight = MatrixProductOperator.random(input_shape, output_shape, bond_shape)
        self.trainable_tensor_weights.append({"name": name, "weight": weight})

    def add_bias(self, size, name=None, initializer="normal"):
        pass


class Linear(Layer):
    # core/model.py:162-206

    def __init__(self, input_units, output_units, core=1, bond=None, bias_initializer="normal",
                 weights_initializer="normal", activation="relu", name="linear"):
        super().__init__(name)
        self.input_units = input_units
        self.output_units = output_units
        self.core = core
        self.bond_dimension = bond
        self.input_core_dim = round(self.input_units ** (1 / self.core))
        self.output_core_dim = round(self.output_units ** (1 / self.core))
        self._input_shape = (self.input_core_dim,) * self.core
        self._output_shape = (self.output_core_dim,) * self.core
        self._bond_shape = (self.bond_dimension,) * (self.core - 1)
        self.bias_initializer = bias_initializer
        self.weights_initializer = weights_initializer
        self.activation = activation
        self.type = "linear"

    def build(self, input_shape):
        if self.weights_initializer == "normal":
            self.add_weight(self._input_shape, self._output_shape, bond_shape=self._bond_shape, name="weight",
                            initializer="normal")
        else:
            self.trainable_tensor_weights.append({"name": "", "weight": self.weights_initializer})

    def __call__(self, inputs):
        super().__call__(inputs)
        weight = self.trainable_tensor_weights[0]["weight"]
        if isinstance(inputs, (list, tuple)):
            return matmul_batch(weight, inputs)
        return weight @ inputs


class Output(Layer):
    # core/model.py:208-218

    def __init__(self, output_shape, name="output"):
        super().__init__(name)
        self.output_shape = output_shape
        self.type = "output"

    def __call__(self, inputs):
        if isinstance(inputs, (list, tuple)):
            return [x.to_tensor().reshape(self.output_shape) for x in inputs]
        return inputs.to_tensor().reshape(self.output_shape)

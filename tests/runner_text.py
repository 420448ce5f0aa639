"""What the reference's Go runner hands to gocca, restated for the tests:
the pooled partition layout, the OKL preamble text, the kernel signature
and the argument list. Test infrastructure only.

- offsets: CalculateAlignedOffsetsAndSize (runner/builder/builder.go:113-159),
  through the oracle's restatement;
- preamble: GeneratePreamble (runner/builder/builder.go:172-410): int_t
  typedef, NPART, KpartMax, static matrices (column-major, %.16e or %.8ff),
  X_PART macros, MATMUL_X / MATMUL_ADD_X macros for static and device
  matrices;
- signature and argument order: K | device matrices sorted by name | arrays
  in configuration order as X_global, X_offsets | scalars
  (runner/kernel_execution.go:150-227, runner/kernel_config.go:186-246);
- device matrices are stored column-major (runner/runner.go:375-415).
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

CT = {np.dtype(np.float64): "double", np.dtype(np.float32): "float",
      np.dtype(np.int32): "int", np.dtype(np.int64): "long"}


class RunnerText:
    def __init__(self, K, int_type: str = "INT64"):
        self.K = [int(k) for k in K]
        self.int_type = int_type
        self.int_np = np.int64 if int_type == "INT64" else np.int32
        self.arrays = {}    # name -> dict(values per element, dtype, host data, output)
        self.statics = {}   # name -> (A, dtype)
        self.devmats = {}   # name -> (A, dtype)
        self.scalars = {}   # name -> dtype

    # ---- bindings
    def array(self, name, per_elem: int, dtype=np.float64, data=None, output=False):
        """A pooled array with per_elem values per element; data: the flat
        concatenation over partitions (element-major within each)."""
        self.arrays[name] = {"per": int(per_elem), "dtype": np.dtype(dtype), "output": output,
                             "data": None if data is None else np.asarray(data, dtype=dtype).reshape(-1)}

    def static_matrix(self, name, A, dtype=np.float64):
        self.statics[name] = (np.asarray(A, dtype=np.float64), np.dtype(dtype))

    def device_matrix(self, name, A, dtype=np.float64):
        self.devmats[name] = (np.asarray(A, dtype=np.float64), np.dtype(dtype))

    def scalar(self, name, dtype=np.float64):
        self.scalars[name] = np.dtype(dtype)

    # ---- layout
    def offsets(self, name):
        a = self.arrays[name]
        nbytes = sum(self.K) * a["per"] * a["dtype"].itemsize
        dt = {4: 3 if a["dtype"].kind == "i" else 1, 8: 4 if a["dtype"].kind == "i" else 2}
        offs, total = O.aligned_offsets(self.K, nbytes, dt[a["dtype"].itemsize], 1)
        return np.asarray(offs, dtype=np.int64), int(total)

    def pooled(self, name):
        """Host image of the pooled allocation (partition p at offsets[p])."""
        a = self.arrays[name]
        offs, total = self.offsets(name)
        out = np.zeros(total // a["dtype"].itemsize, dtype=a["dtype"])
        if a["data"] is not None:
            pos = 0
            for p, k in enumerate(self.K):
                n = k * a["per"]
                out[offs[p]:offs[p] + n] = a["data"][pos:pos + n]
                pos += n
        return out

    def unpool(self, name, image):
        a = self.arrays[name]
        offs, _ = self.offsets(name)
        return np.concatenate([image[offs[p]:offs[p] + k * a["per"]] for p, k in enumerate(self.K)])

    # ---- text
    def preamble(self) -> str:
        lines = [f"typedef {'long' if self.int_type == 'INT64' else 'int'} int_t;", "",
                 f"#define NPART {len(self.K)}", f"#define KpartMax {max(self.K)}", ""]
        if self.statics:
            lines.append("// Static matrices")
            for name, (A, dt) in self.statics.items():
                rows, cols = A.shape
                ty = CT[dt]
                lines.append(f"const {ty} {name}[{cols}][{rows}] = {{")
                for j in range(cols):
                    vals = ", ".join((f"{A[i, j]:.8f}f" if dt == np.float32 else f"{A[i, j]:.16e}")
                                     for i in range(rows))
                    lines.append("  {" + vals + "}" + ("," if j < cols - 1 else ""))
                lines.append("};")
            lines.append("")
        lines.append("// Partition access macros")
        for name in sorted(self.arrays):
            ty = CT[self.arrays[name]["dtype"]]
            lines.append(f"#define {name}_PART(part) (({ty}*)({name}_global + {name}_offsets[part]))")
        lines.append("")
        text = "\n".join(lines) + "\n"
        for name, (A, dt) in list(self.statics.items()) + list(self.devmats.items()):
            text += self._matmul_macro(name, A, dt, name in self.statics)
        return text

    @staticmethod
    def _matmul_macro(name, A, dt, static) -> str:
        rows, cols = A.shape
        ty, zero = ("float", "0.0f") if dt == np.float32 else ("double", "0.0")
        acc = (f"{name}[jj][ii] * (IN)[elem * {cols} + jj]" if static
               else f"{name}[jj * {rows} + ii] * (IN)[elem * {cols} + jj]")
        out = ""
        for macro, assign in ((f"MATMUL_{name}", "="), (f"MATMUL_ADD_{name}", "+=")):
            out += (f"#define {macro}(IN, OUT, K_VAL) {{\\\n"
                    f"  for (int ii = 0; ii < {rows}; ++ii) {{\\\n"
                    "    for (int elem = 0; elem < KpartMax; ++elem; @inner) {\\\n"
                    "      if (elem < (K_VAL)) {\\\n"
                    f"        {ty} sum = {zero};\\\n"
                    f"        for (int jj = 0; jj < {cols}; ++jj) {{\\\n"
                    f"          sum += {acc};\\\n"
                    "        }\\\n"
                    f"        (OUT)[elem * {rows} + ii] {assign} sum;\\\n"
                    "      }\\\n"
                    "    }\\\n"
                    "  }\\\n"
                    "}\n\n")
        return out

    def params(self, order):
        """Kernel parameters for a configuration that names ``order`` (arrays,
        matrices, scalars) -- the argument list of kernel_execution.go:150-227."""
        ps = [("K", "const int_t* K")]
        for name in sorted(n for n in order if n in self.devmats):
            ps.append((name, f"const {CT[self.devmats[name][1]]}* {name}"))
        for name in order:
            if name in self.arrays:
                a = self.arrays[name]
                const = "" if a["output"] else "const "
                ps.append((name + "_global", f"{const}{CT[a['dtype']]}* {name}_global"))
                ps.append((name + "_offsets", f"const int_t* {name}_offsets"))
        for name in order:
            if name in self.scalars:
                ps.append((name, f"{CT[self.scalars[name]]} {name}"))
        return ps

    def signature(self, order) -> str:
        return ",\n\t".join(p for _, p in self.params(order))

    def source(self, kernel_text: str) -> str:
        """BuildKernel: preamble + "\\n" + kernel source (runner/runner.go:88-99)."""
        return self.preamble() + "\n" + kernel_text


class DeviceRun:
    """Device buffers for a RunnerText and the argument list of one launch."""

    def __init__(self, rt: RunnerText, dev):
        from dgkernel_amd import _lib
        self.rt, self.dev, self.L = rt, dev, _lib
        self.buf = {"K": dev.upload(np.asarray(rt.K, dtype=rt.int_np))}
        for name in rt.arrays:
            self.buf[name + "_global"] = dev.upload(rt.pooled(name))
            offs, _ = rt.offsets(name)
            self.buf[name + "_offsets"] = dev.upload(offs.astype(rt.int_np))
        for name, (A, dt) in rt.devmats.items():
            self.buf[name] = dev.upload(np.ascontiguousarray(A.T).astype(dt).reshape(-1))

    def args(self, order, scalars=None):
        scalars = scalars or {}
        out = []
        for key, _ in self.rt.params(order):
            if key in self.rt.scalars:
                kind = {np.dtype(np.float64): self.L.ARG_F64, np.dtype(np.float32): self.L.ARG_F32,
                        np.dtype(np.int32): self.L.ARG_I32,
                        np.dtype(np.int64): self.L.ARG_I64}[self.rt.scalars[key]]
                out.append((kind, scalars[key]))
            else:
                out.append(self.buf[key])
        return out

    def result(self, name):
        a = self.rt.arrays[name]
        img = self.buf[name + "_global"].to_host(dtype=a["dtype"])
        return self.rt.unpool(name, img)

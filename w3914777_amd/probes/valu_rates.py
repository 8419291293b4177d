"""Issue cost of the VALU instructions the snowflake sweeps are built from.

    python -m w3914777_amd.probes.valu_rates

Runs the kernels of valu_rates.hip (built by w3914777_amd.build into
libvalu_rates.so) and prints, per instruction kind, the cycles one wave64
instruction takes on a SIMD at the device's nominal clock: SIMD cycles over
the launch (CUs x 4 SIMDs x elapsed time x clock) divided by the
wave-instructions issued (workgroups x 4 waves x iterations x instructions per
iteration)."""
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvalu_rates.so")

# kernel index -> (name, wave-instructions per loop iteration, values per instruction)
KERNELS = [
    ("v_fma_f32", 8, 1),
    ("v_fma_f64", 8, 1),
    ("v_pk_fma_f16", 8, 2),
    ("v_rcp_f32", 8, 1),
    ("v_rcp_f16 (+1/8 v_pack_b32_f16)", 9, 1),
    ("v_add_f32 + 2 conversions", 24, 1),
    ("v_pk_fma_f16 + 2 v_fma_mix_f32", 12, 1),
]


def main(reps=5):
    if not os.path.exists(LIB):
        raise SystemExit(f"{LIB} is not built (python -m w3914777_amd.build)")
    L = ctypes.CDLL(LIB)
    L.valu_rate_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    iters = L.valu_rate_iters()
    print(f"{'instruction':<34} {'ms':>8} {'cycles/wave-instr':>18}")
    for k, (name, per_iter, _) in enumerate(KERNELS):
        ms, cus, clk = ctypes.c_float(), ctypes.c_int(), ctypes.c_int()
        rc = L.valu_rate_run(k, reps, ctypes.byref(ms), ctypes.byref(cus), ctypes.byref(clk))
        if rc != 0:
            raise SystemExit(f"valu_rate_run({k}) failed: {rc}")
        blocks = cus.value * 8
        winstr = reps * blocks * 4.0 * iters * per_iter
        cycles = cus.value * 4 * ms.value * 1e-3 * clk.value * 1e3
        print(f"{name:<34} {ms.value:>8.3f} {cycles / winstr:>18.2f}")
    print(f"({cus.value} CUs, nominal clock {clk.value / 1000:.0f} MHz, {blocks} workgroups of 256, {reps} launches)")
    return 0


if __name__ == "__main__":
    sys.exit(main())

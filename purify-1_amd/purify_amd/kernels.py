"""Kernel names of cpp/purify/kernels.h:15-18 (kernels::kernel and kernel_from_string)."""
import enum


class kernel(enum.IntEnum):  # noqa: N801 -- mirrors the reference's enum class
    kb = 0
    gauss = 1
    box = 2
    pswf = 3
    kbmin = 4
    gauss_alt = 5

    @staticmethod
    def coerce(k):
        if isinstance(k, str):
            return kernel_from_string[k]
        return kernel(int(k))


kernel_from_string = {"kb": kernel.kb, "gauss": kernel.gauss, "box": kernel.box, "pswf": kernel.pswf,
                      "kbmin": kernel.kbmin, "gauss_alt": kernel.gauss_alt}

"""Command-line information tool, the counterpart of NeMo's ``nemo`` utility
(src/app/nemo.cpp:22-98): ``python -m nemosim_amd [OPTIONS]`` prints the
library version, the HIP devices a simulation can run on, or the plugin
directories; with no option it prints the usage and exits with status 1."""
import argparse
import ctypes
import os
import sys

from . import _lib


def _devices(lib):
    count = ctypes.c_uint(0)
    if lib.nemo_cuda_device_count(ctypes.byref(count)) != 0 or count.value == 0:
        return []
    out = []
    for d in range(count.value):
        descr = ctypes.c_char_p()
        if lib.nemo_cuda_device_description(d, ctypes.byref(descr)) == 0 and descr.value:
            out.append((d, descr.value.decode()))
    return out


def system_plugin_path():
    """The plugins/ directory beside the library (searched last)."""
    return os.path.join(os.path.dirname(os.path.abspath(_lib.LIB_PATH)), "plugins")


def user_plugin_path():
    """$HOME/.nemo/plugins (searched before the system directory)."""
    return os.path.join(os.path.expanduser("~"), ".nemo", "plugins")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m nemosim_amd", add_help=False)
    ap.add_argument("-h", "--help", action="store_true", help="print this message")
    ap.add_argument("-v", "--version", action="store_true", help="print version number")
    ap.add_argument("-l", "--list-devices", action="store_true", help="print the available simulation devices")
    ap.add_argument("--system-plugin-path", action="store_true", help="print the system plugin path")
    ap.add_argument("--user-plugin-path", action="store_true", help="print the user plugin path")
    try:
        args = ap.parse_args(argv)
    except SystemExit:
        return 1
    if args.help:
        ap.print_help()
        return 0
    if args.version:
        print(_lib.load().nemo_version().decode())
        return 0
    if args.list_devices:
        devs = _devices(_lib.load())
        if not devs:
            print("No HIP devices available")
        for d, descr in devs:
            print("%d: %s" % (d, descr))
        return 0
    if args.system_plugin_path:
        print(system_plugin_path())
        return 0
    if args.user_plugin_path:
        print(user_plugin_path())
        return 0
    ap.print_help()
    return 1


if __name__ == "__main__":
    sys.exit(main())

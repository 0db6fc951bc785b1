"""Command-line front end: JSON config in, JSON results out (src/bqa/cli.py:1-175).

    python -m bqa_amd.cli [-i CONFIG.json] [-o RESULT.json] [-l LOG_LEVEL] [-b BACKEND] [-h]

The options and their behaviour are the reference's: ``-i/--input`` and
``-o/--output`` take paths relative to the working directory that must end in
``.json`` (stdin / stdout when absent); ``-l/--log-level`` is one of DEBUG,
INFO, WARNING, ERROR (INFO by default); ``-h/--help`` prints the usage and
exits 0; an unknown argument prints a message and exits 1; any failure prints
the chain of causes ("caused by: ...") to stderr and exits 1.  The result is
the ``run_qa`` list, e.g. ``[["bloch_vectors", [[x, y, z], ...]],
["measurement_outcomes", [1, -1, ...]]]``.

``-b/--backend`` is this package's addition: it overrides the config's
``"backend"`` key, so a config written for the reference (``"numpy"``) runs
here unchanged with ``-b mi355x``.
"""
from __future__ import annotations

import json
import logging
import os
import sys
from pathlib import Path

LOG_LEVELS = ("DEBUG", "INFO", "WARNING", "ERROR")
DEFAULT_LOG_LEVEL = "INFO"


class CliError(Exception):
    """A command-line or file error (src/bqa/cli.py:29-30)."""


def usage(prog: str = "python -m bqa_amd.cli") -> str:
    cwd = os.getcwd()
    return f"""usage: {prog} [options]

    options:
      -i | --input  [PATH]         path to a `*.json` config relative to `{cwd}`; stdin if not given
      -o | --output [PATH]         path to a `*.json` result file relative to `{cwd}`; stdout if not given
      -l | --log-level [LOG_LEVEL] one of {', '.join(LOG_LEVELS)}; {DEFAULT_LOG_LEVEL} if not given
      -b | --backend [NAME]        run on backend NAME whatever the config's "backend" key says
      -h | --help                  show this message and exit"""


def json_path(arg: str) -> Path:
    """A `.json` path resolved against the working directory (src/bqa/cli.py:37-44)."""
    try:
        p = (Path(os.getcwd()) / Path(arg)).resolve()
    except (RuntimeError, OSError) as e:
        raise CliError(f"Error while checking {arg} file") from e
    if p.suffix != ".json":
        raise CliError(f"{p} must have .json suffix")
    return p


def parse_args(args: list) -> dict:
    """Options -> {"input", "output", "log-level", "backend", "help"}; raises CliError on a missing value
    and ValueError (with the reference's message) on an unknown argument."""
    opts = {"input": None, "output": None, "log-level": DEFAULT_LOG_LEVEL, "backend": None, "help": False}
    it = iter(args)
    for arg in it:
        if arg in ("-h", "--help"):
            opts["help"] = True
            return opts
        if arg in ("-i", "--input", "-o", "--output", "-l", "--log-level", "-b", "--backend"):
            val = next(it, None)
            if arg in ("-i", "--input"):
                if val is None:
                    raise CliError(f"Input file is not provided after {arg} key")
                opts["input"] = json_path(val)
            elif arg in ("-o", "--output"):
                if val is None:
                    raise CliError(f"Output file is not provided after {arg} key")
                opts["output"] = json_path(val)
            elif arg in ("-l", "--log-level"):
                if val is None:
                    raise CliError(f"Logging level is not provided after {arg} key, must be one of "
                                   f"{', '.join(LOG_LEVELS)}")
                if val not in LOG_LEVELS:
                    raise CliError(f"{val} cannot be interpeted as a logging level, must be one of "
                                   f"{', '.join(LOG_LEVELS)}")
                opts["log-level"] = val
            else:
                if val is None:
                    raise CliError(f"Backend name is not provided after {arg} key")
                opts["backend"] = val
            continue
        raise ValueError(f"Invalid command line argument {arg}")
    return opts


def read_config(src):
    """The JSON config from a path, or from a text stream (src/bqa/cli.py:111-128)."""
    try:
        if isinstance(src, Path):
            with src.open("r") as f:
                return json.load(f)
        return json.load(src)
    except (json.JSONDecodeError, PermissionError, IsADirectoryError, UnicodeError, OSError) as e:
        raise CliError("Error while parsing json data") from e


def write_result(dst, data) -> None:
    """src/bqa/cli.py:131-148."""
    try:
        if isinstance(dst, Path):
            with dst.open("w") as f:
                json.dump(data, f)
        else:
            json.dump(data, dst)
    except (PermissionError, IsADirectoryError, OSError) as e:
        raise CliError("Error while writing result") from e


def format_error(e: BaseException) -> str:
    """The error and its chain of causes, one per line (src/bqa/cli.py:151-156)."""
    msgs = []
    while e is not None:
        msgs.append(str(e))
        e = e.__cause__
    return "\ncaused by: ".join(msgs)


def main(argv=None, stdin=None, stdout=None, stderr=None) -> int:
    """Run the CLI; returns the exit status (0 or 1)."""
    argv = sys.argv[1:] if argv is None else list(argv)
    stdin = sys.stdin if stdin is None else stdin
    stdout = sys.stdout if stdout is None else stdout
    stderr = sys.stderr if stderr is None else stderr
    try:
        opts = parse_args(argv)
    except ValueError as e:
        print(f"{e}, run `{sys.argv[0]} --help` to get the documentation", file=stdout)
        return 1
    except CliError as e:
        print(format_error(e), file=stderr)
        return 1
    if opts["help"]:
        print(usage(), file=stdout)
        return 0
    try:
        logging.basicConfig(level=getattr(logging, opts["log-level"]),
                            format="%(asctime)s [%(levelname)s] %(message)s", datefmt="%Y-%m-%d %H:%M:%S")
        config = read_config(opts["input"] if opts["input"] is not None else stdin)
        if opts["backend"] is not None and isinstance(config, dict):
            config = dict(config, backend=opts["backend"])
        from .core import run_qa
        result = run_qa(config)
        write_result(opts["output"] if opts["output"] is not None else stdout, result)
    except Exception as e:  # the reference reports every failure the same way (cli.py:166-175)
        print(format_error(e), file=stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())

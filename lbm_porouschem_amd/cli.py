"""Command line entry of the HIP solver, the counterpart of the reference's `lbm`
command (src/LBM/cli.py:8-49):

    python -m lbm_porouschem_amd.cli -c config.yaml [-v 0|1|2] [--precision fp32|fp64|fp16]
                                     [--collision bgk|mrt]

The configuration file is read with a safe YAML loader that also resolves YAML 1.2
floats (2e-5), as the reference's ruamel.yaml does.  If the named file is missing,
the first *.yaml in the working directory is used, as the reference does.
"""
import logging
import os
import sys

import click


@click.command()
@click.option("--config", "-c", default="config.yaml", help="the configuration of simulation case")
@click.option("--verbose", "-v", default=1, help="log level 0:warning;1:info;2:debug")
@click.option("--precision", default="fp32", type=click.Choice(["fp32", "fp64", "fp16"]))
@click.option("--collision", default="bgk", type=click.Choice(["bgk", "mrt"]))
def run(config, verbose, precision, collision):
    levels = {0: logging.WARNING, 1: logging.INFO, 2: logging.DEBUG}
    logger = logging.getLogger("lbm_porouschem_amd")
    if not logger.handlers:
        logger.addHandler(logging.StreamHandler(sys.stdout))
    logger.setLevel(levels.get(verbose, logging.INFO))
    if not os.path.exists(config):
        found = [p for p in sorted(os.listdir(os.curdir)) if p.endswith(".yaml")]
        if not found:
            raise click.UsageError(f"{config} not found and no *.yaml file in the working directory")
        logger.info(f"{config} not found. Found {found[0]} yaml file instead.")
        config = found[0]
    logger.info(f"Loading {config} file")
    from .app import application, load_config
    cfg = load_config(config)
    application(cfg, logger, precision=precision, collision=collision)


if __name__ == "__main__":
    run()

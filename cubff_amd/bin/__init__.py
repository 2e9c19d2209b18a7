"""Built host binaries: `main` (the CLI) and the `cubff` Python module.

Built by cubff_amd/host/Makefile.  Import as `from cubff_amd.bin import cubff`
or put this directory on sys.path and `import cubff`.
"""

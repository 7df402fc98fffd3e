"""INI-style ``.cfg`` reader/writer compatible with oceansar.ocs_io.ConfigFile.

Behaviour follows oceansar/ocs_io/cfg.py:67-198: sections in brackets become
attributes; ';' and '#' start comments (also in-line); values are typed as
int (all digits), bool (True/False), string / bool / float arrays ([...]),
float, else string. The value pattern keeps the reference's character class,
so '+' ends a value (``1e+3`` reads as ``1e``, a string) exactly as there.
"""
import ast
import re

import numpy as np

_COMMENT_OR_BLANK = re.compile(r'(^[;#].*$)|(^[\s\t\r\n]*$)')
_SECTION = re.compile(r'\[(\w+)\]')
_VALUE = re.compile(r"(\w+)\s*=\s*([\s\w\s\\\:\.,/\'\-\[\]]+)")


class Section:
    def __init__(self):
        self.elements = []


def _convert(text):
    if text.isdigit():
        return int(text)
    if text in ('True', 'False'):
        return text == 'True'
    if text[0] == '[':
        if text[1] == "'":
            return text.replace("'", '')[1:-1].split(',')
        if text[1] in ('T', 'F'):
            parts = text[1:-1].split(',')
            return np.array([ast.literal_eval(p.strip()) for p in parts],
                            dtype=bool)
        return np.array(text[1:-1].split(',')).astype(float)
    try:
        return float(text)
    except ValueError:
        return text


class ConfigFile:
    """Parsed configuration; ``cfg.section.key`` access (cfg.py:10-143)."""

    def __init__(self, cfg_file_name, str_mode=False):
        self.cfg_file_name = cfg_file_name
        self.str_mode = str_mode
        self.read()

    def read(self):
        self.sections = [self]
        self.elements = []
        current = self
        with open(self.cfg_file_name, 'r') as fh:
            for line in fh:
                if _COMMENT_OR_BLANK.search(line):
                    continue
                line = line.strip()
                line = re.sub(r'[;#].*$', '', line)
                s = _SECTION.search(line)
                if s:
                    vars(self)[s.group(1)] = Section()
                    self.sections.append(s.group(1))
                    current = vars(self)[s.group(1)]
                    continue
                s = _VALUE.search(line)
                if not s:
                    continue
                key, text = s.group(1), s.group(2)
                vars(current)[key] = text if self.str_mode else _convert(text)
                current.elements.append(key)

    def save(self, alternate_file=None, blank_lines=True):
        """Write back (cfg.py:145-198); comments are not preserved."""
        with open(alternate_file or self.cfg_file_name, 'w') as fh:
            for section in self.sections:
                if section is self:
                    inst = self
                else:
                    fh.write('[%s]\n' % section)
                    inst = vars(self)[section]
                for el in inst.elements:
                    v = vars(inst)[el]
                    if isinstance(v, bool):
                        fh.write('%s=%s\n' % (el, 'True' if v else 'False'))
                    elif isinstance(v, (int, np.integer)):
                        fh.write('%s=%d\n' % (el, v))
                    elif isinstance(v, (float, np.floating)):
                        fh.write('%s=%.8f\n' % (el, v))
                    elif isinstance(v, str):
                        fh.write('%s=%s\n' % (el, v))
                    elif isinstance(v, list):
                        fh.write('%s=[' % el + ','.join("'%s'" % f for f in v) + ']\n')
                    elif isinstance(v, np.ndarray):
                        fh.write('%s=[' % el + ','.join('%.2f' % f for f in v) + ']\n')
                    else:
                        raise TypeError('Variable type not supported')
                if blank_lines:
                    fh.write('\n')

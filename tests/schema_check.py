"""A small JSON Schema checker for the keywords the timing schemas use (type,
const, enum, required, properties, additionalProperties: false, items,
minimum, minLength).  Test helper only: returns a list of error paths."""


def _type_ok(v, t):
    ts = t if isinstance(t, list) else [t]
    for x in ts:
        if x == "null" and v is None:
            return True
        if x == "object" and isinstance(v, dict):
            return True
        if x == "array" and isinstance(v, list):
            return True
        if x == "string" and isinstance(v, str):
            return True
        if x == "boolean" and isinstance(v, bool):
            return True
        if x == "integer" and isinstance(v, int) and not isinstance(v, bool):
            return True
        if x == "number" and isinstance(v, (int, float)) and not isinstance(v, bool):
            return True
    return False


def check(value, schema, path="$"):
    errs = []
    if "type" in schema and not _type_ok(value, schema["type"]):
        return [f"{path}: type {type(value).__name__} not {schema['type']}"]
    if "const" in schema and value != schema["const"]:
        errs.append(f"{path}: {value!r} != {schema['const']!r}")
    if "enum" in schema and value not in schema["enum"]:
        errs.append(f"{path}: {value!r} not in {schema['enum']}")
    if isinstance(value, (int, float)) and not isinstance(value, bool) and "minimum" in schema:
        if value < schema["minimum"]:
            errs.append(f"{path}: {value} < {schema['minimum']}")
    if isinstance(value, str) and "minLength" in schema and len(value) < schema["minLength"]:
        errs.append(f"{path}: shorter than {schema['minLength']}")
    if isinstance(value, dict):
        props = schema.get("properties", {})
        for k in schema.get("required", []):
            if k not in value:
                errs.append(f"{path}: missing {k}")
        for k, v in value.items():
            if k in props:
                errs += check(v, props[k], f"{path}.{k}")
            elif schema.get("additionalProperties") is False:
                errs.append(f"{path}: unexpected {k}")
    if isinstance(value, list) and "items" in schema:
        for i, v in enumerate(value):
            errs += check(v, schema["items"], f"{path}[{i}]")
    return errs
